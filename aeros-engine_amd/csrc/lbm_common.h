// Shared declarations of the MI355X LBM library: host-side setup results and
// the parameter blocks handed to the HIP kernels by value.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <vector>

#include "../../include/aeros_lbm.h"

namespace aeros {

constexpr int Q = 19;   // D3Q19
constexpr int Q7 = 7;   // D3Q7 (first seven D3Q19 directions)

// Per-cell link mask (uint32), built once per mask:
//   bits 0..18  source of direction q is solid/ground (half-way bounce-back);
//               for shell cells the bit refers to the in-domain source after a
//               periodic wrap or a free-slip reflection
//   bit  19     the cell itself is solid or ground (never updated)
//   bits 20..22 Van Driest table index: wall distance 0..5, 6 = no damping
//   bit  23     Smagorinsky strain is defined here (interior, not on the inlet
//               or outlet plane, no solid 6-neighbour)
//   bits 24..29 the cell lies on the domain face x = 0, x = nx-1, y = 0,
//               y = ny-1, z = 0, z = nz-1 (any of them: a shell cell; the
//               inlet and outlet planes are two of these faces)
//   bit  30     pull = inlet equilibrium     bit 31  pull = convective outlet
constexpr uint32_t M_SOLID = 1u << 19;
constexpr int M_VD_SHIFT = 20;
constexpr uint32_t M_VD_MASK = 7u << 20;
constexpr uint32_t M_STRAIN = 1u << 23;
constexpr int M_FACE_SHIFT = 24;
constexpr uint32_t M_X0 = 1u << 24, M_X1 = 1u << 25, M_Y0 = 1u << 26, M_Y1 = 1u << 27, M_Z0 = 1u << 28,
                   M_Z1 = 1u << 29;
constexpr uint32_t M_FACES = 63u << 24;
constexpr uint32_t M_PULL_INLET = 1u << 30;
constexpr uint32_t M_PULL_OUTLET = 1u << 31;
constexpr uint32_t M_LINKS = (1u << 19) - 1u;
// face bit of the plane where axis a (0..2) is 0 (hi = 0) or N-1 (hi = 1)
constexpr uint32_t face_bit(int axis, int hi) { return 1u << (M_FACE_SHIFT + 2 * axis + hi); }

// Result of the host-side part of initLBM (domain, mask, wall distance, tau).
struct Setup {
    aeros_lbm_config cfg{};
    int nx = 0, ny = 0, nz = 0;
    int64_t total = 0;
    float min[3] = {0, 0, 0};
    float cell[3] = {0.1f, 0.1f, 0.1f};
    std::vector<uint8_t> solid, ground, wall;   // per cell
    int64_t fluid = 1;
    float inlet[3] = {0, 0, 0};                 // computeInletVelocityLB
    int axis = 0, sign = 1;                     // computeFlowAxis
    float tau = 0.6f, viscosity = 0.033f, U0 = 0.1f;
    float re_lb = 0, re_phys = 0, blockage = 0;
    // thermal (thermalInit)
    bool thermal_ready = false;
    int ylo = 1, yhi = 0;
    float H = 1, tauT = 0.7f, alpha = 0, gravity = 0, rayleigh = 0, prandtl = 0;
    bool curl_noise = false;                    // inletTurbulence > 0.001
};

// Free-slip faces must be known faces off the periodic axes, and cannot be
// combined with the thermal lattice: AEROS_OK or AEROS_ERR_ARG with a reason.
int check_boundaries(const aeros_lbm_config& cfg, char* msg, int msglen);
// initLBM's host half: returns AEROS_OK or an error code; msg gets the reason.
int setup_domain(const aeros_lbm_config& cfg, Setup& s, char* msg, int msglen);
// The ground slab of cfg (use_ground / ground_effect) on top of the solid mask:
// non-solid cells up to the ground plane become ground (s.ground).
void apply_ground(Setup& s);
// Everything after the mask is known: wall distance, fluid count, tau, thermal.
int setup_after_mask(Setup& s, char* msg, int msglen);
void config_defaults(aeros_lbm_config* c);
void inlet_velocity(const aeros_lbm_config& c, float& ux, float& uy, float& uz);
void flow_axis(const aeros_lbm_config& c, int& axis, int& sign);
// The free stream in world units (flowParams.vx/vy/vz).
void free_stream(const aeros_lbm_config& c, float v[3]);
// Initial temperature field of thermalResetField (host, glibc cos/sin).
void thermal_initial_field(const Setup& s, std::vector<float>& T);
// Initial per-cell velocity with the curl-noise inflow (host, glibc cos).
void curl_noise_velocity(const Setup& s, std::vector<float>& u3);

// Voxelizer (BVH three-ray vote), host side.
int voxelize_host(const float* tris, int64_t ntris, int nx, int ny, int nz, const float mn[3],
                  const float cs[3], int32_t* out);

// The voxelizer's BVH flattened for the device: nodes in build order (node 0
// is the root; a leaf has count > 0 and owns triangles [first, first+count)),
// triangles in leaf order.
struct VoxNode {
    float lo[3], hi[3];
    int32_t left, right, first, count;
};
struct VoxTri {
    float a[3], b[3], c[3];
};
// Drops non-finite and degenerate triangles and builds the BVH exactly as
// voxelize_host does.  Empty `nodes` means no usable triangle.
void voxel_bvh(const float* tris, int64_t ntris, std::vector<VoxNode>& nodes, std::vector<VoxTri>& out_tris);
// Ray jitter of the three-axis vote for a grid of cell size cs.
void voxel_jitter(const float cs[3], float jit[3]);

// STL (host): format of a buffer, and its triangles (9 floats each) with
// per-vertex normals; returns AEROS_OK or AEROS_ERR_ARG.
void stl_detect(const unsigned char* d, size_t size, aeros_stl_info* out);
int stl_parse(const unsigned char* d, size_t size, std::vector<float>& verts, std::vector<float>& normals,
              int64_t* dropped);

// ----------------------------------------------------------------- kernels
template <class R> struct StepParams {
    int nx, ny, nz;
    int64_t n;
    int axis, sign;          // flow axis / sign
    int les, trt, thermal, periodic_x, periodic_z, conv_outlet;
    int first;               // state holds pre-collision populations: collide only
    int write_fields;        // write rho/u/tke/strain fields this step
    int dry;                 // re-run of an aborted step: no abort check, no reduction
    int k;                   // index of this step since the lattice was (re)initialised
    int pending;             // step k - 1 is still to be folded into the status
    R tau0, cs, vel_scale2;  // sanitised tau, Smagorinsky constant, (flowSpeed/U0)^2
    R inv_tau0, s_minus0;    // 1/tau0 and the TRT s- at tau0 (exactly what the kernel would compute)
    R vd[7];                 // Van Driest table
    R fx, fy, fz;            // body force
    R in_u[3];               // inlet velocity
    R feq_in[Q];             // computeEquilibriumFast(q, 1, u_in)
    R conv_c, conv_k2;       // convective outlet c and 0.15 c
    int64_t out_d1, out_d2;  // cell offsets to the planes N-2, N-3 behind the outlet
    // thermal
    R t_inv_tau, t_ref, t_g, t_hi, t_lo;
    R t_hot, t_cold;
    int t_fixed;
    // domain faces: the inlet and outlet planes, the faces a periodic axis
    // wraps and the free-slip faces (shell cells on either of the last two
    // take the general pull)
    uint32_t inlet_face, outlet_face, wrap_faces, slip_faces;
    // cell -> (x, y, z): floor(c / d) = (uint)(c inv_d + half_inv_d) in FP64,
    // exact for c < 2^26 and d < 2^25
    double inv_nx, half_inv_nx, inv_ny, half_inv_ny;
    // reduction
    int64_t fluid_cells;
    int64_t nan_limit;       // total / 10
    float vel_scale;         // flowSpeed / U0 (clamped), for max |u| world
    float conv_threshold;
    // 1 when tau, the body force and the buoyancy bound every collision
    // output of populations in [-1, 10] far below FLT_MAX, so that the
    // non-finite fix-up can be skipped for such cells
    int bounded;
    // the shift of the stored populations of the equilibrium-shifted 16-bit
    // storage: the inflow equilibrium feq_in (w_q for the other storages)
    R st_shift[Q];
};

}  // namespace aeros
