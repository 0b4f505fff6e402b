// ============================================================================
// TEST INFRASTRUCTURE ONLY — CPU oracle for the Aeros LBM hot path.
//
// A restatement, in our own code, of the reference solver's algorithm:
// D3Q19 collide + pull stream (BGK/TRT, Guo forcing, Smagorinsky LES with Van
// Driest damping, half-way bounce-back, periodic / copy-in-place shell (and
// free-slip walls, which the reference lacks),
// equilibrium inlet, convective outlet), diagnostics, divergence reset and the
// end-of-call post-pass; the D3Q7 thermal lattice with Boussinesq buoyancy and
// Nusselt number; the control-volume drag and projected area; and the BVH
// three-ray voxelizer.  Every function cites the reference file:line it
// follows.  Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
// may load this library, and only as the checker; the product never does.
//
// The class is a template on the arithmetic type R.  R = float reproduces the
// reference's float arithmetic operation by operation (no FMA contraction:
// build with -ffp-contract=off); R = double is the FP64 oracle.  The float
// instantiation can also quantise every stored population to IEEE half or
// bfloat16 deviations f - w (round to nearest even), which is what the 16-bit
// storage modes of the GPU path store.
//
// The layout is the reference's AoS f[cell*19 + q]; loops mirror the
// reference's interior + shell split so that the GPU's unified per-cell
// formulation is checked against an independent one.
// ============================================================================
#include "../include/aeros_lbm.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- tables
// D3Q19 velocity set and opposites in the reference's order
// (aeros/src/lbm.cpp:92-109).
constexpr int Q = 19;
const int CV[Q][3] = {
    {0,0,0},
    {1,0,0}, {-1,0,0}, {0,1,0}, {0,-1,0}, {0,0,1}, {0,0,-1},
    {1,1,0}, {-1,-1,0}, {1,-1,0}, {-1,1,0},
    {1,0,1}, {-1,0,-1}, {1,0,-1}, {-1,0,1},
    {0,1,1}, {0,-1,-1}, {0,1,-1}, {0,-1,1}};
const int OPP[Q] = {0, 2,1,4,3,6,5, 8,7,10,9, 12,11,14,13, 16,15,18,17};

// Van Driest factor 1 - exp(-d/2), d = 0..6, as tabulated in
// aeros/src/lbm.cpp:112-120.
const double VD_TABLE[7] = {0.0, 0.3934693403, 0.6321205588, 0.7768698399,
                            0.8646647168, 0.9179150014, 0.9502129316};
static_assert((float)0.3934693403 == 0.3934693403f, "constant rounding");
static_assert((float)0.9502129316 == 0.9502129316f, "constant rounding");
static_assert((float)0.09 == 0.09f, "constant rounding");
static_assert((float)0.15 == 0.15f, "constant rounding");
static_assert((float)0.1875 == 0.1875f, "constant rounding");
static_assert((float)1e-6 == 1e-6f, "constant rounding");

// D3Q7 (aeros/src/thermal_lbm.cpp:31-41): the first seven D3Q19 directions.
constexpr int Q7 = 7;

template <class R> struct K {
    static R w(int q) {   // aeros/src/lbm.cpp:100-105: 1/3, 1/18, 1/36 computed in R
        return q == 0 ? R(1) / R(3) : (q <= 6 ? R(1) / R(18) : R(1) / R(36));
    }
    static R w7(int q) { return q == 0 ? R(0.25) : R(0.125); }
    static R cs2() { return R(1) / R(3); }    // aeros/src/lbm.cpp:121
    static R vd(int d) { return R(VD_TABLE[d]); }
};

// ---------------------------------------------------------------- 16-bit storage
// Round-to-nearest-even quantisers used by the 16-bit storage modes.
float q_half(float x) {
    uint32_t u; std::memcpy(&u, &x, 4);
    const uint32_t sign = u & 0x80000000u;
    uint32_t a = u & 0x7fffffffu;
    uint16_t h;
    if (a >= 0x7f800000u) {                       // inf / nan
        h = (uint16_t)((a > 0x7f800000u) ? 0x7e00 : 0x7c00);
    } else if (a >= 0x477ff000u) {                // rounds to >= 65536: overflow
        h = 0x7c00;
    } else if (a < 0x38800000u) {                 // below the smallest normal half
        // subnormal half: value = m * 2^-24
        float f; std::memcpy(&f, &a, 4);
        const float scaled = f * 16777216.0f;     // exact (power of two)
        float r = std::nearbyint(scaled);         // default rounding mode: nearest even
        h = (uint16_t)r;
    } else {
        const uint32_t mant = a & 0x7fffffu;
        const int exp = (int)(a >> 23) - 127 + 15;
        uint32_t hm = mant >> 13;
        const uint32_t rem = mant & 0x1fffu;
        uint32_t hv = ((uint32_t)exp << 10) | hm;
        if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) hv += 1;
        h = (uint16_t)hv;
    }
    // decode back to float
    const uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
    float out;
    if (e == 0) out = (float)m * (1.0f / 16777216.0f);
    else if (e == 31) out = m ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
    else { uint32_t bits = ((e - 15 + 127) << 23) | (m << 13); std::memcpy(&out, &bits, 4); }
    return sign ? -out : out;
}

float q_bf16(float x) {
    uint32_t u; std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;   // nan
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    float out; std::memcpy(&out, &u, 4);
    return out;
}

// ---------------------------------------------------------------- params
struct Params {
    aeros_lbm_config c;
};

bool finite(float x) { return std::isfinite(x); }
bool finite(double x) { return std::isfinite(x); }

// airSutherlandViscosity (aeros/src/flow_field.h:46-54)
float sutherland(float Tk) {
    float T = Tk;
    if (!std::isfinite(T) || T < 50.0f) T = 288.15f;
    if (T > 2000.0f) T = 2000.0f;
    const float T0 = 273.15f, S = 110.4f, mu0 = 1.716e-5f;
    float mu = mu0 * std::pow(T / T0, 1.5f) * (T0 + S) / (T + S);
    if (!std::isfinite(mu) || mu < 1e-7f) mu = 1.81e-5f;
    return mu;
}

// glm::radians in float
float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

// computeInletVelocityLB (aeros/src/lbm.cpp:186-199)
void inlet_velocity(const aeros_lbm_config& c, float& ux, float& uy, float& uz) {
    const float az = radians(c.flow_azimuth_deg), el = radians(c.flow_elevation_deg);
    const float ce = cosf(el), se = sinf(el), ca = cosf(az), sa = sinf(az);
    float dx = ce * ca, dy = se, dz = ce * sa;
    float len = sqrtf(dx * dx + dy * dy + dz * dz);
    if (len < 1e-6f || !std::isfinite(len)) { dx = 1; dy = 0; dz = 0; len = 1; }
    dx /= len; dy /= len; dz /= len;
    float U0 = c.U0;
    if (!std::isfinite(U0) || U0 < 1e-6f) U0 = 0.08f;
    if (U0 > 0.15f) U0 = 0.15f;
    ux = dx * U0; uy = dy * U0; uz = dz * U0;
}

// computeFlowAxis (aeros/src/lbm.cpp:201-214); glm::length = sqrt(dot), dot = (x*x+y*y)+z*z
void flow_axis(const aeros_lbm_config& c, int& axis, int& sign, float dir[3]) {
    const float az = radians(c.flow_azimuth_deg), el = radians(c.flow_elevation_deg);
    const float ce = cosf(el), se = sinf(el), ca = cosf(az), sa = sinf(az);
    float d[3] = {ce * ca, se, ce * sa};
    const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len < 1e-6f || !std::isfinite(len)) { d[0] = 1; d[1] = 0; d[2] = 0; }
    else { d[0] /= len; d[1] /= len; d[2] /= len; }
    const float ax = std::fabs(d[0]), ay = std::fabs(d[1]), azv = std::fabs(d[2]);
    if (ax >= ay && ax >= azv) { axis = 0; sign = (d[0] > 0) ? 1 : -1; }
    else if (ay >= azv) { axis = 1; sign = (d[1] > 0) ? 1 : -1; }
    else { axis = 2; sign = (d[2] > 0) ? 1 : -1; }
    dir[0] = d[0]; dir[1] = d[1]; dir[2] = d[2];
}

// ---------------------------------------------------------------- equilibria
// equilibriumKernel (aeros/src/lbm.cpp:141-144): w*rho*(1 + 3cu + 4.5cu^2 - 1.5u^2)
template <class R> R eq_kernel(int q, R rho, R ux, R uy, R uz, R usqr) {
    const R cu = (R)CV[q][0] * ux + (R)CV[q][1] * uy + (R)CV[q][2] * uz;
    return K<R>::w(q) * rho * (R(1) + cu * R(3) + cu * cu * R(4.5) - usqr * R(1.5));
}

// computeEquilibriumFast (aeros/src/lbm.cpp:146-184): guarded form
template <class R> R eq_guarded(int q, R rho, R ux, R uy, R uz, R usqr) {
    if (!finite(rho)) rho = R(1);
    if (usqr > R(0.09)) {
        R s = std::sqrt(R(0.09) / usqr);
        ux *= s; uy *= s; uz *= s;
        usqr = ux * ux + uy * uy + uz * uz;
    }
    R cu = (R)CV[q][0] * ux + (R)CV[q][1] * uy + (R)CV[q][2] * uz;
    if (!finite(cu)) cu = R(0);
    R feq = K<R>::w(q) * rho * (R(1) + cu * R(3) + cu * cu * R(4.5) - usqr * R(1.5));
    if (!finite(feq)) feq = K<R>::w(q) * rho;
    return feq;
}

// ---------------------------------------------------------------- BVH voxelizer
// Restatement of meshbvh (aeros/src/mesh_bvh.h): median split on the longest
// centroid axis, leaves of 4, depth <= 48; double Moller-Trumbore with a
// relative determinant epsilon; three jittered axis rays, majority vote.
struct V3 { float x, y, z; float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); } };
struct Tri { V3 v0, v1, v2, lo, hi, cen; };

V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

// aeros/src/mesh_bvh.h:62-96
bool ray_tri(const V3& o, const V3& d, const Tri& t, double minT, double maxT) {
    const double e1x = (double)t.v1.x - t.v0.x, e1y = (double)t.v1.y - t.v0.y, e1z = (double)t.v1.z - t.v0.z;
    const double e2x = (double)t.v2.x - t.v0.x, e2y = (double)t.v2.y - t.v0.y, e2z = (double)t.v2.z - t.v0.z;
    const double dx = d.x, dy = d.y, dz = d.z;
    const double px = dy * e2z - dz * e2y, py = dz * e2x - dx * e2z, pz = dx * e2y - dy * e2x;
    const double det = e1x * px + e1y * py + e1z * pz;
    const double scale = std::fabs(e1x * e2y - e1y * e2x) + std::fabs(e1y * e2z - e1z * e2y) +
                         std::fabs(e1z * e2x - e1x * e2z) + 1e-30;
    if (std::fabs(det) < 1e-12 * scale) return false;
    const double inv = 1.0 / det;
    const double tx = (double)o.x - t.v0.x, ty = (double)o.y - t.v0.y, tz = (double)o.z - t.v0.z;
    const double u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const double qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    const double v = (dx * qx + dy * qy + dz * qz) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    const double tt = (e2x * qx + e2y * qy + e2z * qz) * inv;
    if (!(tt > minT) || tt >= maxT) return false;
    return true;
}

struct Bvh {
    struct Node { V3 lo{0,0,0}, hi{0,0,0}; int left = -1, right = -1, start = 0, count = 0; };
    std::vector<Node> nodes;
    std::vector<Tri> tris;

    int build(int start, int count, int depth) {
        const int idx = (int)nodes.size();
        nodes.push_back(Node{});
        V3 lo = tris[start].lo, hi = tris[start].hi, cmin = tris[start].cen, cmax = tris[start].cen;
        for (int i = start + 1; i < start + count; ++i) {
            lo = vmin(lo, tris[i].lo); hi = vmax(hi, tris[i].hi);
            cmin = vmin(cmin, tris[i].cen); cmax = vmax(cmax, tris[i].cen);
        }
        nodes[idx].lo = lo; nodes[idx].hi = hi;
        if (count <= 4 || depth >= 48) { nodes[idx].start = start; nodes[idx].count = count; return idx; }
        const V3 ext{cmax.x - cmin.x, cmax.y - cmin.y, cmax.z - cmin.z};
        int axis = 0;
        if (ext.y > ext.x) axis = 1;
        if (ext.z > ext[axis]) axis = 2;
        if (!(ext[axis] > 1e-12f)) { nodes[idx].start = start; nodes[idx].count = count; return idx; }
        const int mid = count / 2;
        std::nth_element(tris.begin() + start, tris.begin() + start + mid, tris.begin() + start + count,
                         [axis](const Tri& a, const Tri& b) { return a.cen[axis] < b.cen[axis]; });
        const int l = build(start, mid, depth + 1);
        const int r = build(start + mid, count - mid, depth + 1);
        nodes[idx].left = l; nodes[idx].right = r; nodes[idx].count = 0;
        return idx;
    }

    // aeros/src/mesh_bvh.h:241-256
    static bool slab(const V3& o, const float inv[3], const Node& n, float& tn0, float& tf0) {
        float t0 = -std::numeric_limits<float>::infinity(), t1 = std::numeric_limits<float>::infinity();
        for (int a = 0; a < 3; ++a) {
            const float oa = o[a], ia = inv[a], lo = n.lo[a], hi = n.hi[a];
            float tn = (lo - oa) * ia, tf = (hi - oa) * ia;
            if (tn > tf) std::swap(tn, tf);
            if (std::isnan(tn) || std::isnan(tf)) continue;
            if (tn > t0) t0 = tn;
            if (tf < t1) t1 = tf;
            if (t1 < t0) return false;
        }
        tn0 = t0; tf0 = t1;
        return t1 >= 0.0f;
    }

    // aeros/src/mesh_bvh.h:116-163: the hit count does not depend on the
    // traversal order, so a plain stack suffices.
    int count(const V3& o, const V3& d) const {
        if (nodes.empty()) return 0;
        const float inf = std::numeric_limits<float>::infinity();
        const float inv[3] = {d.x != 0.0f ? 1.0f / d.x : inf, d.y != 0.0f ? 1.0f / d.y : inf,
                              d.z != 0.0f ? 1.0f / d.z : inf};
        int st[128]; int sp = 0; int hits = 0;
        float a, b;
        if (!slab(o, inv, nodes[0], a, b)) return 0;
        st[sp++] = 0;
        while (sp > 0) {
            const Node& n = nodes[st[--sp]];
            if (n.count > 0) {
                for (int i = 0; i < n.count; ++i)
                    if (ray_tri(o, d, tris[n.start + i], 1e-7, 1e30)) ++hits;
                continue;
            }
            if (slab(o, inv, nodes[n.left], a, b)) st[sp++] = n.left;
            if (slab(o, inv, nodes[n.right], a, b)) st[sp++] = n.right;
        }
        return hits;
    }

    // aeros/src/mesh_bvh.h:168-185
    bool inside(const V3& p, const V3& cell) const {
        if (nodes.empty()) return false;
        const V3 j{0.731f * 1e-3f * cell.x, 0.377f * 1e-3f * cell.y, 0.557f * 1e-3f * cell.z};
        int votes = 0;
        if (count({p.x, p.y + j.y, p.z + j.z}, {1, 0, 0}) & 1) ++votes;
        if (count({p.x + j.z, p.y, p.z + j.x}, {0, 1, 0}) & 1) ++votes;
        if (count({p.x + j.y, p.y + j.x, p.z}, {0, 0, 1}) & 1) ++votes;
        return votes >= 2;
    }
};

// ---------------------------------------------------------------- solver
template <class R> class Sim {
public:
    aeros_lbm_config cfg{};
    int storage = AEROS_FP32;     // FP32/FP64 exact, FP16/BF16 quantised (R = float), FP32_FP64C
                                  // deviations f - w and the velocity stencil rounded to float (R = double)
    int Nx = 0, Ny = 0, Nz = 0, total = 0;
    aeros_lbm_config user_cfg{};   // the parameters init was called with (tau and U0 unsanitised)
    float minX = 0, minY = 0, minZ = 0, csX = 0.1f, csY = 0.1f, csZ = 0.1f;
    std::vector<R> f, fNext;
    std::vector<R> rho, ux, uy, uz, tke, strain, uxW, uyW, uzW, pres;
    std::vector<char> solid, ground;
    std::vector<uint8_t> wallDist;
    int fluidCells = 1;
    float tau = 0.6f, viscosity = 0.033f, U0 = 0.1f;
    float reLB = 0, rePhys = 0, blockage = 0;
    // scalars
    int curStep = 0; bool converged = false; int resets = 0;
    float lastDivergedRho = 0.0f; int lastDivergedStep = -1;
    float avgRho = 1, avgKE = 0, maxVelLB = 0, maxVelW = 0, tkeMean = 0, conv = 0;
    long long negPops = 0, badPops = 0, rhoOut = 0, rhoNonFin = 0, uLim = 0, pBad = 0;
    float prevKinetic = 0;
    float pMin = 0, pMax = 0, resRho = 0, resU = 0, prevMaxVelLB = 0;
    std::vector<float> prevRho; int prevTotal = -1, prevStep = -1;
    // thermal (aeros/src/thermal_lbm.cpp)
    bool thermalReady = false;
    std::vector<char> tsolid;
    std::vector<R> g, gNext, T, buoy;
    int yLo = 1, yHi = 0; float H = 1, tauT = 0.7f;
    float tNu = 0, tAlpha = 0, tRa = 0, tPr = 0, tG = 0, tNusselt = 0, tMeanT = 0.5f;

    // inflow equilibrium f_q^eq(1, u_in): the shift of the equilibrium-shifted
    // 16-bit storages (no counterpart in the reference)
    R eqs[Q] = {};
    bool quantised() const {
        return storage == AEROS_FP16 || storage == AEROS_BF16 || storage == AEROS_FP32_FP64C ||
               storage == AEROS_FP16_EQ || storage == AEROS_BF16_EQ;
    }
    R qz(R x, int q) const {   // storage quantisation of a population
        if (storage == AEROS_FP32_FP64C) return K<R>::w(q) + (R)(float)(x - K<R>::w(q));
        if (storage == AEROS_FP16) return K<R>::w(q) + (R)q_half((float)(x - K<R>::w(q)));
        if (storage == AEROS_BF16) return K<R>::w(q) + (R)q_bf16((float)(x - K<R>::w(q)));
        // stored half(float(f - e) * 2^12), read h * 2^-12 + e (one float rounding each)
        if (storage == AEROS_FP16_EQ)
            return eqs[q] + (R)(q_half((float)(x - eqs[q]) * 4096.0f) * (1.0f / 4096.0f));
        if (storage == AEROS_BF16_EQ) return eqs[q] + (R)q_bf16((float)(x - eqs[q]));
        return x;
    }
    R qu(R x) const {          // storage quantisation of the velocity stencil input
        if (storage == AEROS_FP32_FP64C) return (R)(float)x;
        if (storage == AEROS_FP16 || storage == AEROS_BF16 || storage == AEROS_FP16_EQ || storage == AEROS_BF16_EQ)
            return (R)q_half((float)x);
        return x;
    }
    void quantize_cell(std::vector<R>& a, int cell) {
        if (!quantised()) return;
        for (int q = 0; q < Q; ++q) a[(size_t)cell * Q + q] = qz(a[(size_t)cell * Q + q], q);
    }

    int idx(int x, int y, int z) const { return (z * Ny + y) * Nx + x; }

    // ------------------------------------------------------------ initLBM
    // aeros/src/lbm.cpp:405-886
    int init(const aeros_lbm_config& c) {
        cfg = c;
        user_cfg = c;
        const int32_t* vox = c.voxels;
        const bool haveVox = vox && c.vox_n[0] > 0 && c.vox_n[1] > 0 && c.vox_n[2] > 0;
        const long long capPadded = c.cell_cap > 0 ? c.cell_cap : 8LL * 1024 * 1024;
        const long long capHard = c.cell_cap > 0 ? c.cell_cap : 20LL * 1024 * 1024;
        if (!haveVox) {   // aeros/src/lbm.cpp:422-429
            Nx = Ny = Nz = c.voxel_resolution;
            minX = minY = minZ = -1.0f;
        } else {           // aeros/src/lbm.cpp:430-514
            const int bN[3] = {c.vox_n[0], c.vox_n[1], c.vox_n[2]};
            float cs[3];
            for (int i = 0; i < 3; ++i) {
                cs[i] = (c.vox_max[i] - c.vox_min[i]) / std::max(1, bN[i]);
                if (!std::isfinite(cs[i]) || cs[i] < 1e-6f) cs[i] = 0.1f;
            }
            float Lref = std::max(std::max(c.vox_max[0] - c.vox_min[0], c.vox_max[1] - c.vox_min[1]),
                                  c.vox_max[2] - c.vox_min[2]);
            if (!std::isfinite(Lref) || Lref < 1e-6f) Lref = 1.0f;
            int ax, sg; float fd[3];
            flow_axis(c, ax, sg, fd);
            float padUser = c.domain_padding;
            if (!std::isfinite(padUser)) padUser = 1.0f;
            padUser = std::min(std::max(padUser, 0.0f), 1.0f);
            float pMinv[3], pMaxv[3];
            for (int i = 0; i < 3; ++i) {
                const float d = std::isfinite(fd[i]) ? fd[i] : 0.0f;
                const float up = 2.0f * Lref * std::fabs(d) * padUser;
                const float down = 4.0f * Lref * std::fabs(d) * padUser;
                const float lat = 2.5f * Lref * (1.0f - std::fabs(d)) * padUser;
                pMinv[i] = std::max(d > 0.0f ? up : down, lat);
                pMaxv[i] = std::max(d > 0.0f ? down : up, lat);
            }
            auto totalFor = [&](float scale) {
                long long t = 1;
                for (int i = 0; i < 3; ++i) {
                    int extra = (int)std::lround((pMinv[i] + pMaxv[i]) * scale / cs[i]);
                    if (extra < 0) extra = 0;
                    t *= (long long)(bN[i] + extra);
                }
                return t;
            };
            float padScale = 1.0f;
            while (padScale > 0.05f && totalFor(padScale) > capPadded) padScale *= 0.75f;
            if (totalFor(padScale) > capPadded) padScale = 0.0f;
            minX = c.vox_min[0] - pMinv[0] * padScale;
            minY = c.vox_min[1] - pMinv[1] * padScale;
            minZ = c.vox_min[2] - pMinv[2] * padScale;
            int nOut[3];
            for (int i = 0; i < 3; ++i) {
                int eMin = (int)std::ceil((pMinv[i] * padScale - 1e-9f) / cs[i]);
                int eMax = (int)std::ceil((pMaxv[i] * padScale - 1e-9f) / cs[i]);
                if (padScale <= 0.0f) { eMin = eMax = 0; }
                nOut[i] = bN[i] + std::max(0, eMin) + std::max(0, eMax);
                if (nOut[i] < 8) nOut[i] = 8;
            }
            Nx = nOut[0]; Ny = nOut[1]; Nz = nOut[2];
            csX = cs[0]; csY = cs[1]; csZ = cs[2];
        }
        const long long tot = (long long)Nx * Ny * Nz;
        if (tot > capPadded) return AEROS_ERR_TOO_LARGE;          // aeros/src/lbm.cpp:526-545
        if (!(csX > 1e-6f) || !std::isfinite(csX)) csX = 2.0f / std::max(1, Nx);
        if (!(csY > 1e-6f) || !std::isfinite(csY)) csY = 2.0f / std::max(1, Ny);
        if (!(csZ > 1e-6f) || !std::isfinite(csZ)) csZ = 2.0f / std::max(1, Nz);
        if (tot <= 0 || tot > capHard) return AEROS_ERR_TOO_LARGE;  // aeros/src/lbm.cpp:557-564
        total = (int)tot;
        try {
            rho.assign(total, R(1)); ux.assign(total, R(0)); uy.assign(total, R(0)); uz.assign(total, R(0));
            uxW.assign(total, R(0)); uyW.assign(total, R(0)); uzW.assign(total, R(0)); pres.assign(total, R(0));
            tke.assign(total, R(0)); strain.assign(total, R(0));
            solid.assign(total, 0); ground.assign(total, 0);
            f.assign((size_t)total * Q, R(0)); fNext.assign((size_t)total * Q, R(0));
        } catch (const std::bad_alloc&) { return AEROS_ERR_NOMEM; }

        // mask by cell-centre world coordinates (aeros/src/lbm.cpp:596-622)
        if (haveVox) {
            const float bx = (c.vox_max[0] - c.vox_min[0]) / (float)c.vox_n[0];
            const float by = (c.vox_max[1] - c.vox_min[1]) / (float)c.vox_n[1];
            const float bz = (c.vox_max[2] - c.vox_min[2]) / (float)c.vox_n[2];
            const long long vt = (long long)c.vox_n[0] * c.vox_n[1] * c.vox_n[2];
            for (int z = 0; z < Nz; ++z) {
                const float pz = minZ + (z + 0.5f) * csZ;
                const int gz = (int)std::floor((pz - c.vox_min[2]) / (bz > 1e-9f ? bz : 1e-9f));
                for (int y = 0; y < Ny; ++y) {
                    const float py = minY + (y + 0.5f) * csY;
                    const int gy = (int)std::floor((py - c.vox_min[1]) / (by > 1e-9f ? by : 1e-9f));
                    for (int x = 0; x < Nx; ++x) {
                        const float px = minX + (x + 0.5f) * csX;
                        const int gx = (int)std::floor((px - c.vox_min[0]) / (bx > 1e-9f ? bx : 1e-9f));
                        if (gx < 0 || gx >= c.vox_n[0] || gy < 0 || gy >= c.vox_n[1] || gz < 0 || gz >= c.vox_n[2]) continue;
                        const long long gi = ((long long)gz * c.vox_n[1] + gy) * c.vox_n[0] + gx;
                        if (gi < 0 || gi >= vt) continue;
                        solid[idx(x, y, z)] = (vox[gi] == 1) ? 1 : 0;
                    }
                }
            }
        }
        ground_slab();
        return finish_init();
    }

    // ground slab over the solid mask (aeros/src/lbm.cpp:624-643)
    void ground_slab() {
        const aeros_lbm_config& c = cfg;
        if (!(c.ground_effect || c.use_ground)) return;
        float groundY = c.vox_min[1] + c.ground_effect_height;
        if (c.use_ground) groundY = c.vox_min[1] + c.ground_height;
        if (!std::isfinite(groundY)) groundY = c.vox_min[1];
        int gy0 = (int)((groundY - minY) / csY);
        if (gy0 < 0) gy0 = 0;
        if (gy0 >= Ny) gy0 = Ny - 1;
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y <= gy0; ++y)
                for (int x = 0; x < Nx; ++x) { const int cc = idx(x, y, z); if (!solid[cc]) ground[cc] = 1; }
    }

    // Everything after the mask transfer: wall distance, fluid count, the
    // equilibrium initialisation, tau and the thermal module.  Also the path
    // taken when the caller hands in a lattice-resolution mask.
    int finish_init() {
        const aeros_lbm_config& c = cfg;
        // wall distance: two-pass 6-neighbour chamfer, saturating (aeros/src/lbm.cpp:647-675)
        wallDist.assign(total, 255);
        for (int cc = 0; cc < total; ++cc) if (solid[cc] || ground[cc]) wallDist[cc] = 0;
        const int sy = Nx, sz = Nx * Ny;
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y < Ny; ++y) {
                const int row = (z * Ny + y) * Nx;
                for (int x = 0; x < Nx; ++x) {
                    unsigned d = wallDist[row + x];
                    if (d == 0) continue;
                    if (x > 0 && wallDist[row + x - 1] + 1u < d) d = wallDist[row + x - 1] + 1u;
                    if (y > 0 && wallDist[row - sy + x] + 1u < d) d = wallDist[row - sy + x] + 1u;
                    if (z > 0 && wallDist[row - sz + x] + 1u < d) d = wallDist[row - sz + x] + 1u;
                    wallDist[row + x] = (uint8_t)d;
                }
            }
        for (int z = Nz - 1; z >= 0; --z)
            for (int y = Ny - 1; y >= 0; --y) {
                const int row = (z * Ny + y) * Nx;
                for (int x = Nx - 1; x >= 0; --x) {
                    unsigned d = wallDist[row + x];
                    if (d == 0) continue;
                    if (x < Nx - 1 && wallDist[row + x + 1] + 1u < d) d = wallDist[row + x + 1] + 1u;
                    if (y < Ny - 1 && wallDist[row + sy + x] + 1u < d) d = wallDist[row + sy + x] + 1u;
                    if (z < Nz - 1 && wallDist[row + sz + x] + 1u < d) d = wallDist[row + sz + x] + 1u;
                    wallDist[row + x] = (uint8_t)d;
                }
            }
        fluidCells = 0;
        for (int cc = 0; cc < total; ++cc) if (!solid[cc] && !ground[cc]) ++fluidCells;
        if (fluidCells < 1) fluidCells = 1;

        // equilibrium initialisation (aeros/src/lbm.cpp:682-739)
        float inUx, inUy, inUz;
        inlet_velocity(c, inUx, inUy, inUz);
        for (int cc = 0; cc < total; ++cc) {
            float fu[3] = {0, 0, 0};
            if (!solid[cc] && !ground[cc]) { fu[0] = inUx; fu[1] = inUy; fu[2] = inUz; }
            if (c.inlet_turbulence > 0.001f && !solid[cc] && !ground[cc]) {
                // curl-noise (aeros/src/lbm.cpp:698-726); parity unpinned
                const float turb = c.inlet_turbulence * 0.04f;
                const int x = cc % Nx, y = (cc / Nx) % Ny, z = cc / (Nx * Ny);
                const float fx = (float)x / Nx * 6.28f, fy = (float)y / Ny * 6.28f, fz = (float)z / Nz * 6.28f;
                const float dAzdy = cosf(fx * 1.2f + fy * 0.8f) * 0.8f;
                const float dAydz = cosf(fz * 1.1f + fx * 0.9f) * 1.1f;
                const float dAxdz = cosf(fy * 1.3f + fz * 0.7f) * 0.7f;
                const float dAzdx = cosf(fx * 1.2f + fy * 0.8f) * 1.2f;
                const float dAydx = cosf(fz * 1.1f + fx * 0.9f) * 0.9f;
                const float dAxdy = cosf(fy * 1.3f + fz * 0.7f) * 1.3f;
                const float rx = (dAzdy - dAydz) * turb, ry = (dAxdz - dAzdx) * turb, rz = (dAydx - dAxdy) * turb;
                const float fLen = sqrtf(inUx * inUx + inUy * inUy + inUz * inUz);
                if (fLen > 1e-6f) {
                    const float d0 = inUx / fLen, d1 = inUy / fLen, d2 = inUz / fLen;
                    const float along = rx * d0 + ry * d1 + rz * d2;
                    fu[0] += rx - d0 * along; fu[1] += ry - d1 * along; fu[2] += rz - d2 * along;
                }
            }
            R u0 = fu[0], u1 = fu[1], u2 = fu[2];
            R us = u0 * u0 + u1 * u1 + u2 * u2;
            if (us > R(0.09)) {
                const R s = std::sqrt(R(0.09) / us);
                u0 *= s; u1 *= s; u2 *= s; us = u0 * u0 + u1 * u1 + u2 * u2;
            }
            for (int q = 0; q < Q; ++q) f[(size_t)cc * Q + q] = eq_guarded<R>(q, R(1), u0, u1, u2, us);
        }
        std::fill(fNext.begin(), fNext.end(), R(0));
        {   // the inflow equilibrium exactly as the inlet pull forms it
            const R u0 = inUx, u1 = inUy, u2 = inUz, us = u0 * u0 + u1 * u1 + u2 * u2;
            for (int q = 0; q < Q; ++q) eqs[q] = eq_guarded<R>(q, R(1), u0, u1, u2, us);
        }
        if (quantised())
            for (int cc = 0; cc < total; ++cc) { quantize_cell(f, cc); quantize_cell(fNext, cc); }

        // tau from Re or from the user (aeros/src/lbm.cpp:741-815)
        const float mu = sutherland(c.air_temperature);
        float L = c.max_dim; if (L < 1e-6f) L = 1.0f;
        float rhoA = c.air_density; if (!std::isfinite(rhoA) || rhoA < 0.01f) rhoA = 1.225f;
        float U0t = std::sqrt(inUx * inUx + inUy * inUy + inUz * inUz);
        if (!std::isfinite(U0t) || U0t < 1e-6f) U0t = 0.08f;
        float RePhys = rhoA * c.flow_speed * L / mu;
        if (!std::isfinite(RePhys) || RePhys < 1.0f) RePhys = 1e4f;
        float csAvg = (csX + csY + csZ) / 3.0f;
        if (csAvg < 1e-6f) csAvg = L / Nx;
        float L_LB = L / csAvg;
        if (L_LB < 5.0f) L_LB = 5.0f;
        if (L_LB > 200.0f) L_LB = 200.0f;
        const float cs2f = 1.0f / 3.0f;
        const float nuRe = (U0t * L_LB) / RePhys;
        float tauPhys = nuRe / cs2f + 0.5f;
        if (!std::isfinite(tauPhys)) tauPhys = 1.5f;
        float tf;
        if (c.auto_viscosity) tf = std::min(std::max(tauPhys, 0.51f), 1.5f);
        else { tf = c.tau; if (!std::isfinite(tf)) tf = 0.6f; tf = std::min(std::max(tf, 0.51f), 1.5f); }
        tau = tf;
        viscosity = (tau - 0.5f) * cs2f;
        U0 = U0t;
        cfg.tau = tau; cfg.U0 = U0t;       // initLBM writes lbmParams.tau / U0 back
        {
            const float nuA = (tf - 0.5f) * cs2f;
            reLB = (nuA > 1e-12f) ? (U0t * L_LB / nuA) : 0.0f;
            if (!std::isfinite(reLB)) reLB = 0.0f;
        }
        rePhys = RePhys;
        blockage = 0.0f;
        if (c.frontal_area > 0.0f) {       // aeros/src/lbm.cpp:826-853
            int ax, sg; float fd[3];
            flow_axis(c, ax, sg, fd);
            float sec = ax == 0 ? (Ny * csY) * (Nz * csZ) : (ax == 1 ? (Nx * csX) * (Nz * csZ) : (Nx * csX) * (Ny * csY));
            if (std::isfinite(sec) && sec > 1e-9f && std::isfinite(c.frontal_area)) blockage = c.frontal_area / sec;
        }
        curStep = 0; converged = false;
        rho.assign(total, R(1)); ux.assign(total, R(0)); uy.assign(total, R(0)); uz.assign(total, R(0));
        uxW.assign(total, R(0)); uyW.assign(total, R(0)); uzW.assign(total, R(0)); pres.assign(total, R(0));
        tke.assign(total, R(0)); strain.assign(total, R(0));
        if (c.thermal) {   // aeros/src/lbm.cpp:876-885
            std::vector<char> mask(total);
            for (int cc = 0; cc < total; ++cc) mask[cc] = (solid[cc] || ground[cc]) ? 1 : 0;
            if (!thermal_init(mask, viscosity)) cfg.thermal = 0;
        }
        return AEROS_OK;
    }

    // A new solid mask on the same grid: the ground slab goes back on top of
    // it, as initLBM applies it after every mask transfer.
    int set_mask(const uint8_t* m) {
        cfg = user_cfg;   // initLBM starts again from the caller's parameters
        for (int cc = 0; cc < total; ++cc) { solid[cc] = m[cc] ? 1 : 0; ground[cc] = 0; }
        ground_slab();
        return finish_init();
    }

    // ------------------------------------------------------------ thermal
    // thermalInit (aeros/src/thermal_lbm.cpp:84-140)
    bool thermal_init(const std::vector<char>& mask, float nuLB) {
        thermalReady = false;
        if (Nx < 4 || Ny < 4 || Nz < 1) return false;
        if ((long long)Nx * Ny * Nz > 20LL * 1024 * 1024 && cfg.cell_cap <= 0) return false;
        if (!std::isfinite(nuLB) || nuLB <= 0.0f) nuLB = 0.05f;
        tsolid = mask;
        yLo = -1; yHi = -1;
        for (int y = 0; y < Ny; ++y) {
            bool fl = false;
            for (int z = 0; z < Nz && !fl; ++z)
                for (int x = 0; x < Nx && !fl; ++x)
                    if (!tsolid[idx(x, y, z)]) fl = true;
            if (fl) { if (yLo < 0) yLo = y; yHi = y; }
        }
        if (yLo < 0 || yHi <= yLo) return false;
        H = (float)(yHi - yLo + 1);
        const float nu = nuLB;
        tPr = std::max(0.05f, cfg.prandtl);
        tAlpha = nu / tPr;
        tauT = 0.5f + tAlpha / 0.25f;
        if (!std::isfinite(tauT) || tauT < 0.505f) tauT = 0.505f;
        if (tauT > 1.5f) { tauT = 1.5f; tAlpha = 0.25f * (tauT - 0.5f); tPr = nu / tAlpha; }
        const float dT = cfg.hot_temp - cfg.cold_temp;
        tG = cfg.rayleigh * nu * tAlpha / (std::max(1e-6f, std::fabs(dT)) * H * H * H);
        tRa = tG * std::fabs(dT) * H * H * H / (nu * tAlpha);
        T.assign(total, R(0.5f * (cfg.hot_temp + cfg.cold_temp)));
        buoy.assign(total, R(0));
        g.assign((size_t)total * Q7, R(0)); gNext.assign((size_t)total * Q7, R(0));
        thermalReady = true;
        thermal_reset();
        return true;
    }

    bool wall_temp(int y, float* Tw) const {   // aeros/src/thermal_lbm.cpp:60-65
        if (!cfg.fixed_wall_temps) return false;
        if (y <= 0) { *Tw = cfg.hot_temp; return true; }
        if (y >= Ny - 1) { *Tw = cfg.cold_temp; return true; }
        return false;
    }

    R geq(int i, R t, R vx, R vy, R vz) const {   // aeros/src/thermal_lbm.cpp:53-56
        const R cu = (R)CV[i][0] * vx + (R)CV[i][1] * vy + (R)CV[i][2] * vz;
        return K<R>::w7(i) * t * (R(1) + R(4) * cu);
    }

    void thermal_diag() {   // computeDiagnostics (aeros/src/thermal_lbm.cpp:67-80)
        double s = 0.0;
        for (int z = 0; z < Nz; ++z)
            for (int y = yLo; y <= yHi; ++y)
                for (int x = 0; x < Nx; ++x) { const int cc = idx(x, y, z); if (!tsolid[cc]) s += T[cc]; }
        const long n = (long)Nx * (long)Nz * (long)(yHi - yLo + 1);
        tMeanT = n > 0 ? (float)(s / (double)n) : 0.5f;
        tRa = cfg.rayleigh;
        tPr = cfg.prandtl;
    }

    // thermalResetField (aeros/src/thermal_lbm.cpp:157-186)
    void thermal_reset() {
        if (!thermalReady) return;
        const float Th = cfg.hot_temp, Tc = cfg.cold_temp, dT = Th - Tc, amp = cfg.perturbation * dT;
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y < Ny; ++y)
                for (int x = 0; x < Nx; ++x) {
                    const int cc = idx(x, y, z);
                    float t;
                    if (tsolid[cc]) { float Tw; t = wall_temp(y, &Tw) ? Tw : 0.5f * (Th + Tc); }
                    else {
                        const float s = (float)(y - yLo) / std::max(1.0f, H - 1.0f);
                        t = Th + (Tc - Th) * s;
                        const float px = 6.2831853f * (float)x / (float)Nx;
                        const float py = 3.1415926f * s;
                        t += amp * std::cos(px) * std::sin(py);
                    }
                    T[cc] = (R)t;
                }
        for (int cc = 0; cc < total; ++cc)
            for (int i = 0; i < Q7; ++i) g[(size_t)cc * Q7 + i] = geq(i, T[cc], R(0), R(0), R(0));
        thermal_diag();
    }

    // thermalUpload (aeros/src/thermal_lbm.cpp:188-197)
    bool thermal_upload(const float* t) {
        if (!thermalReady) return false;
        for (int cc = 0; cc < total; ++cc) T[cc] = (R)t[cc];
        for (int cc = 0; cc < total; ++cc)
            for (int i = 0; i < Q7; ++i) g[(size_t)cc * Q7 + i] = geq(i, T[cc], R(0), R(0), R(0));
        thermal_diag();
        return true;
    }

    // thermalStep (aeros/src/thermal_lbm.cpp:199-355): pull-stream (interior,
    // then the shell), then BGK, buoyancy and the Nusselt number.
    void thermal_step() {
        if (!thermalReady || !cfg.thermal) return;
        const R invTau = R(1) / (R)tauT;
        int off[Q7];
        for (int i = 0; i < Q7; ++i) off[i] = -(CV[i][2] * Ny * Nx) - CV[i][1] * Nx - CV[i][0];
        const int yA = cfg.fixed_wall_temps ? 2 : 1;
        const int yB = cfg.fixed_wall_temps ? Ny - 2 : Ny - 1;
        #pragma omp parallel for
        for (int z = 1; z < Nz - 1; ++z)
            for (int y = yA; y < yB; ++y)
                for (int x = 1; x < Nx - 1; ++x) {
                    const int cc = idx(x, y, z);
                    if (tsolid[cc]) { for (int i = 0; i < Q7; ++i) gNext[(size_t)cc * Q7 + i] = g[(size_t)cc * Q7 + i]; continue; }
                    for (int i = 0; i < Q7; ++i) {
                        const int src = cc + off[i];
                        gNext[(size_t)cc * Q7 + i] = tsolid[src] ? g[(size_t)cc * Q7 + OPP[i]] : g[(size_t)src * Q7 + i];
                    }
                }
        #pragma omp parallel for
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y < Ny; ++y) {
                const bool yEdge = (y < yA || y >= yB), zEdge = (z == 0 || z == Nz - 1);
                for (int x = 0; x < Nx; ++x) {
                    if (!yEdge && !zEdge && x > 0 && x < Nx - 1) continue;
                    const int cc = idx(x, y, z);
                    if (tsolid[cc]) { for (int i = 0; i < Q7; ++i) gNext[(size_t)cc * Q7 + i] = g[(size_t)cc * Q7 + i]; continue; }
                    for (int i = 0; i < Q7; ++i) {
                        int sx = x - CV[i][0], sy = y - CV[i][1], sz = z - CV[i][2];
                        const bool outX = (sx < 0 || sx >= Nx), outY = (sy < 0 || sy >= Ny), outZ = (sz < 0 || sz >= Nz);
                        if (outX) sx = (sx < 0) ? Nx - 1 : 0;
                        if (outZ) sz = (sz < 0) ? Nz - 1 : 0;
                        float Tw = 0.0f; bool dir = false, bounce = false;
                        if (outY) { bounce = true; dir = wall_temp(sy < 0 ? 0 : Ny - 1, &Tw); }
                        else if (tsolid[idx(sx, sy, sz)]) { bounce = true; dir = wall_temp(sy, &Tw); }
                        if (bounce) {
                            const R gOpp = g[(size_t)cc * Q7 + OPP[i]];
                            gNext[(size_t)cc * Q7 + i] = dir ? (-gOpp + R(2) * K<R>::w7(i) * (R)Tw) : gOpp;
                        } else gNext[(size_t)cc * Q7 + i] = g[(size_t)idx(sx, sy, sz) * Q7 + i];
                    }
                }
            }
        const R Tref = (R)(0.5f * (cfg.hot_temp + cfg.cold_temp));
        const R gLb = (R)tG;
        const R hiT = (R)(cfg.hot_temp + 1.0f), loT = (R)(cfg.cold_temp - 1.0f);
        #pragma omp parallel for
        for (int cc = 0; cc < total; ++cc) {
            if (tsolid[cc]) { buoy[cc] = R(0); continue; }
            const R vx = ux[cc], vy = uy[cc], vz = uz[cc];
            R gi[Q7]; R t = R(0);
            for (int i = 0; i < Q7; ++i) { gi[i] = gNext[(size_t)cc * Q7 + i]; t += gi[i]; }
            if (!finite(t)) t = Tref;
            t = std::min(hiT, std::max(loT, t));
            for (int i = 0; i < Q7; ++i) {
                const R e = geq(i, t, vx, vy, vz);
                gNext[(size_t)cc * Q7 + i] = gi[i] - (gi[i] - e) * invTau;
            }
            T[cc] = t;
            buoy[cc] = gLb * (t - Tref);
        }
        double qs = 0.0; long qn = 0;
        for (int z = 0; z < Nz; ++z)
            for (int x = 0; x < Nx; ++x) {
                const int cc = idx(x, yLo, z);
                if (tsolid[cc]) continue;
                qs += (double)((R)cfg.hot_temp - T[cc]) / 0.5;
                ++qn;
            }
        if (qn > 0) {
            const float dT = std::max(1e-6f, std::fabs(cfg.hot_temp - cfg.cold_temp));
            tNusselt = (float)(qs / (double)qn) * H / dT;
        }
        thermal_diag();
        g.swap(gNext);
    }

    // ------------------------------------------------------------ step
    bool is_inlet(int ax, int sg, int x, int y, int z) const {
        if (ax == 0) return sg > 0 ? x == 0 : x == Nx - 1;
        if (ax == 1) return sg > 0 ? y == 0 : y == Ny - 1;
        return sg > 0 ? z == 0 : z == Nz - 1;
    }
    bool is_outlet(int ax, int sg, int x, int y, int z) const {
        if (ax == 0) return sg > 0 ? x == Nx - 1 : x == 0;
        if (ax == 1) return sg > 0 ? y == Ny - 1 : y == 0;
        return sg > 0 ? z == Nz - 1 : z == 0;
    }

    // stepLBMCPU (aeros/src/lbm.cpp:907-1709)
    void step(int steps) {
        if (total <= 0) return;
        float tau0 = cfg.tau;
        if (!std::isfinite(tau0) || tau0 < 0.51f) tau0 = 0.55f;
        if (tau0 > 1.5f) tau0 = 1.5f;
        float inUx, inUy, inUz;
        inlet_velocity(cfg, inUx, inUy, inUz);
        float U0mag = std::sqrt(inUx * inUx + inUy * inUy + inUz * inUz);
        if (!std::isfinite(U0mag) || U0mag < 1e-6f) U0mag = 0.08f;
        int ax, sg; float fd[3];
        flow_axis(cfg, ax, sg, fd);
        float velScale = cfg.flow_speed / U0mag;
        if (!std::isfinite(velScale) || velScale > 1000.0f) velScale = 20.0f;
        if (velScale < 0.01f) velScale = 0.01f;
        const R velScale2 = (R)velScale * (R)velScale;
        float Cs = cfg.smagorinsky_c;
        if (!std::isfinite(Cs) || Cs < 0.0f) Cs = 0.12f;
        if (Cs > 0.3f) Cs = 0.3f;
        const R cs2 = K<R>::cs2();
        const bool thermalOn = cfg.thermal && thermalReady;
        bool diverged = false;

        for (int s = 0; s < steps; ++s) {
            // strain of the previous step's velocity (aeros/src/lbm.cpp:971-1036)
            if (cfg.use_les) {
                std::vector<R> qux, quy, quz;
                const std::vector<R>* pu[3] = {&ux, &uy, &uz};
                if (quantised()) {
                    qux.resize(total); quy.resize(total); quz.resize(total);
                    for (int cc = 0; cc < total; ++cc) { qux[cc] = qu(ux[cc]); quy[cc] = qu(uy[cc]); quz[cc] = qu(uz[cc]); }
                    pu[0] = &qux; pu[1] = &quy; pu[2] = &quz;
                }
                const std::vector<R>& UX = *pu[0]; const std::vector<R>& UY = *pu[1]; const std::vector<R>& UZ = *pu[2];
                const int sY = Nx, sZ = Nx * Ny;
                #pragma omp parallel for
                for (int z = 0; z < Nz; ++z)
                    for (int y = 0; y < Ny; ++y)
                        for (int x = 0; x < Nx; ++x) {
                            const int cc = idx(x, y, z);
                            if (solid[cc] || ground[cc]) { strain[cc] = R(0); continue; }
                            if (is_inlet(ax, sg, x, y, z) || is_outlet(ax, sg, x, y, z) ||
                                !(x > 0 && x < Nx - 1 && y > 0 && y < Ny - 1 && z > 0 && z < Nz - 1)) { strain[cc] = R(0); continue; }
                            const int xm = cc - 1, xp = cc + 1, ym = cc - sY, yp = cc + sY, zm = cc - sZ, zp = cc + sZ;
                            if (solid[xm] || solid[xp] || solid[ym] || solid[yp] || solid[zm] || solid[zp]) { strain[cc] = R(0); continue; }
                            const R h = R(0.5);
                            const R dux_dx = (UX[xp] - UX[xm]) * h, duy_dy = (UY[yp] - UY[ym]) * h, duz_dz = (UZ[zp] - UZ[zm]) * h;
                            const R dux_dy = (UX[yp] - UX[ym]) * h, dux_dz = (UX[zp] - UX[zm]) * h;
                            const R duy_dx = (UY[xp] - UY[xm]) * h, duy_dz = (UY[zp] - UY[zm]) * h;
                            const R duz_dx = (UZ[xp] - UZ[xm]) * h, duz_dy = (UZ[yp] - UZ[ym]) * h;
                            const R Sxy = h * (dux_dy + duy_dx), Sxz = h * (dux_dz + duz_dx), Syz = h * (duy_dz + duz_dy);
                            const R S2 = dux_dx * dux_dx + duy_dy * duy_dy + duz_dz * duz_dz + R(2) * (Sxy * Sxy + Sxz * Sxz + Syz * Syz);
                            strain[cc] = std::sqrt(R(2) * S2);
                        }
            }

            // collide in place (aeros/src/lbm.cpp:1038-1311); reductions per z-plane, then in order
            std::vector<double> pRho(Nz, 0.0), pKE(Nz, 0.0), pTke(Nz, 0.0);
            std::vector<float> pMaxV(Nz, 0.0f);
            std::vector<long long> pNan(Nz, 0), pNeg(Nz, 0), pBadv(Nz, 0), pRout(Nz, 0), pRnf(Nz, 0), pUl(Nz, 0);
            #pragma omp parallel for
            for (int z = 0; z < Nz; ++z) {
                double aR = 0, aK = 0, aT = 0; float mV = 0;
                long long nNan = 0, nNeg = 0, nBad = 0, nRo = 0, nRn = 0, nUl = 0;
                for (int y = 0; y < Ny; ++y)
                    for (int x = 0; x < Nx; ++x) {
                        const int cc = idx(x, y, z);
                        if (solid[cc] || ground[cc]) {
                            rho[cc] = R(1); ux[cc] = uy[cc] = uz[cc] = R(0); tke[cc] = R(0); strain[cc] = R(0);
                            continue;
                        }
                        R* fc = &f[(size_t)cc * Q];
                        R r = 0, vx = 0, vy = 0, vz = 0;
                        for (int q = 0; q < Q; ++q) {
                            R fi = fc[q];
                            if (!finite(fi)) { fi = K<R>::w(q); nNan++; nBad++; }
                            if (fi < R(-1) || fi > R(10)) { fi = K<R>::w(q); nBad++; }
                            r += fi; vx += fi * (R)CV[q][0]; vy += fi * (R)CV[q][1]; vz += fi * (R)CV[q][2];
                        }
                        if (!finite(r)) { nRn++; nNan++; r = R(1); }
                        else if (r < R(0.5) || r > R(1.5)) { nRo++; nNan++; r = R(1); }
                        const R invRho = R(1) / r;
                        vx *= invRho; vy *= invRho; vz *= invRho;
                        if (!finite(vx)) { vx = vy = vz = R(0); nNan++; }
                        {
                            const R ut = vx * vx + vy * vy + vz * vz;
                            if (ut > R(0.09)) { const R sc = std::sqrt(R(0.09) / ut); vx *= sc; vy *= sc; vz *= sc; }
                        }
                        const bool inl = is_inlet(ax, sg, x, y, z), outl = is_outlet(ax, sg, x, y, z);
                        if (inl && !cfg.periodic_x) { r = R(1); vx = (R)inUx; vy = (R)inUy; vz = (R)inUz; }
                        R tauEff = (R)tau0, sm = R(0);
                        if (cfg.use_les && !inl && !outl) {
                            sm = strain[cc];
                            if (finite(sm) && sm < R(10) && sm > R(1e-6)) {
                                R csr = (R)Cs;
                                const int wd = (int)wallDist[cc];
                                if (wd <= 5) csr *= K<R>::vd(wd);
                                const R nut = csr * csr * sm;
                                tauEff = (R)tau0 + nut / cs2;
                                if (tauEff > R(1.5)) tauEff = R(1.5);
                                if (tauEff < R(0.51)) tauEff = R(0.51);
                            }
                        }
                        R us = vx * vx + vy * vy + vz * vz;
                        if (us > R(0.09)) {
                            nUl++;
                            const R sc = std::sqrt(R(0.09) / us);
                            vx *= sc; vy *= sc; vz *= sc; us = vx * vx + vy * vy + vz * vz;
                        }
                        const R invTau = R(1) / tauEff;
                        const R Fx = (R)cfg.body_force[0];
                        R Fy = (R)cfg.body_force[1];
                        const R Fz = (R)cfg.body_force[2];
                        if (thermalOn && !buoy.empty()) Fy += buoy[cc];
                        const bool hasF = (Fx != R(0) || Fy != R(0) || Fz != R(0));
                        R ex = vx, ey = vy, ez = vz, es = us;
                        if (hasF) {
                            const R hf = R(0.5) / r;
                            ex += Fx * hf; ey += Fy * hf; ez += Fz * hf;
                            es = ex * ex + ey * ey + ez * ez;
                        }
                        R sMinus = invTau;
                        if (cfg.use_trt) {
                            const R m = invTau - R(0.5);
                            const R den = (std::fabs(m) > R(1e-6)) ? (R(0.1875) / m + R(0.5)) : R(1);
                            sMinus = R(1) / std::max(R(0.05), den);
                        }
                        for (int q = 0; q < Q; ++q) {
                            const int qo = OPP[q];
                            if (qo < q) continue;
                            const R feq = eq_kernel<R>(q, r, ex, ey, ez, es);
                            R fi = fc[q];
                            if (!finite(fi)) fi = feq;
                            R srcI = 0;
                            if (hasF) {
                                const R cx = (R)CV[q][0], cy = (R)CV[q][1], cz = (R)CV[q][2];
                                const R cu = cx * ex + cy * ey + cz * ez;
                                srcI = R(3) * K<R>::w(q) * ((cx - ex + R(3) * cu * cx) * Fx + (cy - ey + R(3) * cu * cy) * Fy +
                                                            (cz - ez + R(3) * cu * cz) * Fz);
                            }
                            if (qo == q) {
                                R fn = fi - (fi - feq) * invTau + (R(1) - R(0.5) * invTau) * srcI;
                                if (!finite(fn)) fn = feq;
                                if (fn < R(0)) nNeg++;
                                fc[q] = fn;
                                continue;
                            }
                            const R feqO = eq_kernel<R>(qo, r, ex, ey, ez, es);
                            R fo = fc[qo];
                            if (!finite(fo)) fo = feqO;
                            R srcO = 0;
                            if (hasF) {
                                const R cx = (R)CV[qo][0], cy = (R)CV[qo][1], cz = (R)CV[qo][2];
                                const R cu = cx * ex + cy * ey + cz * ez;
                                srcO = R(3) * K<R>::w(qo) * ((cx - ex + R(3) * cu * cx) * Fx + (cy - ey + R(3) * cu * cy) * Fy +
                                                             (cz - ez + R(3) * cu * cz) * Fz);
                            }
                            const R srcP = R(0.5) * (srcI + srcO), srcM = R(0.5) * (srcI - srcO);
                            const R fp = R(0.5) * (fi + fo), fm = R(0.5) * (fi - fo);
                            const R ep = R(0.5) * (feq + feqO), em = R(0.5) * (feq - feqO);
                            const R fpN = fp - invTau * (fp - ep);
                            const R fmN = fm - sMinus * (fm - em);
                            R nI = fpN + fmN + (R(1) - R(0.5) * invTau) * srcP + (R(1) - R(0.5) * sMinus) * srcM;
                            R nO = fpN - fmN + (R(1) - R(0.5) * invTau) * srcP - (R(1) - R(0.5) * sMinus) * srcM;
                            if (!finite(nI)) nI = feq;
                            if (!finite(nO)) nO = feqO;
                            if (nI < R(0)) nNeg++;
                            if (nO < R(0)) nNeg++;
                            fc[q] = nI; fc[qo] = nO;
                        }
                        quantize_cell(f, cc);
                        rho[cc] = r;
                        if (hasF) {
                            const R hi = R(0.5) / r;
                            ux[cc] = vx + Fx * hi; uy[cc] = vy + Fy * hi; uz[cc] = vz + Fz * hi;
                        } else { ux[cc] = vx; uy[cc] = vy; uz[cc] = vz; }
                        strain[cc] = sm;
                        const R tl = ((R)Cs * sm) * ((R)Cs * sm);
                        R tk = tl * velScale2;
                        if (!finite(tk) || tk < R(0)) tk = R(0);
                        if (tk > R(1.0e6)) tk = R(1.0e6);
                        tke[cc] = tk;
                        const float vm = (float)std::sqrt(us);
                        if (vm > mV) mV = vm;
                        aR += (double)r; aK += (double)us; aT += (double)tk;
                    }
                pRho[z] = aR; pKE[z] = aK; pTke[z] = aT; pMaxV[z] = mV;
                pNan[z] = nNan; pNeg[z] = nNeg; pBadv[z] = nBad; pRout[z] = nRo; pRnf[z] = nRn; pUl[z] = nUl;
            }
            double aR = 0, aK = 0, aT = 0; float mV = 0; long long nNan = 0, nNeg = 0, nBad = 0, nRo = 0, nRn = 0, nUl = 0;
            for (int z = 0; z < Nz; ++z) {
                aR += pRho[z]; aK += pKE[z]; aT += pTke[z]; mV = std::max(mV, pMaxV[z]);
                nNan += pNan[z]; nNeg += pNeg[z]; nBad += pBadv[z]; nRo += pRout[z]; nRn += pRnf[z]; nUl += pUl[z];
            }
            if (nNan > total / 10) { diverged = true; break; }   // aeros/src/lbm.cpp:1313-1319

            stream_step(ax, sg, inUx, inUy, inUz, U0mag);
            f.swap(fNext);
            if (thermalOn) thermal_step();

            // scalars (aeros/src/lbm.cpp:1554-1580)
            curStep++;
            avgRho = (float)(aR / fluidCells);
            avgKE = (float)(aK / fluidCells);
            maxVelLB = mV;
            tkeMean = (float)(aT / fluidCells);
            float scale = cfg.flow_speed / U0mag;
            if (!std::isfinite(scale) || scale > 1000.0f) scale = 20.0f;
            if (scale < 0.01f) scale = 0.01f;
            maxVelW = mV * scale;
            negPops = nNeg; badPops = nBad; rhoOut = nRo; rhoNonFin = nRn; uLim = nUl;
            conv = std::fabs(avgKE - prevKinetic);
            prevKinetic = avgKE;
            converged = conv < cfg.convergence_threshold;
            if (!std::isfinite(avgRho) || avgRho > 1.2f || avgRho < 0.8f) { diverged = true; lastDivergedRho = avgRho; lastDivergedStep = curStep; break; }
        }
        if (diverged) reset_flow(inUx, inUy, inUz);
        post_pass();
    }

    // pull stream with bounce-back, shell, inlet and convective outlet
    // (aeros/src/lbm.cpp:1321-1544)
    void stream_step(int ax, int sg, float inUx, float inUy, float inUz, float U0mag) {
        int off[Q];
        for (int q = 0; q < Q; ++q) off[q] = -(CV[q][2] * Ny * Nx) - CV[q][1] * Nx - CV[q][0];
        #pragma omp parallel for
        for (int z = 1; z < Nz - 1; ++z)
            for (int y = 1; y < Ny - 1; ++y)
                for (int x = 1; x < Nx - 1; ++x) {
                    const int d = idx(x, y, z);
                    if (solid[d] || ground[d]) continue;
                    for (int q = 0; q < Q; ++q) {
                        const int s = d + off[q];
                        fNext[(size_t)d * Q + q] = (solid[s] || ground[s]) ? f[(size_t)d * Q + OPP[q]] : f[(size_t)s * Q + q];
                    }
                }
        #pragma omp parallel for
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y < Ny; ++y) {
                const bool yE = (y == 0 || y == Ny - 1), zE = (z == 0 || z == Nz - 1);
                for (int x = 0; x < Nx; ++x) {
                    if (!yE && !zE && x > 0 && x < Nx - 1) continue;
                    const int d = idx(x, y, z);
                    if (solid[d] || ground[d]) continue;
                    for (int q = 0; q < Q; ++q) {
                        int sx = x - CV[q][0], sy = y - CV[q][1], sz = z - CV[q][2];
                        // free-slip wall half a cell outside a face (no reference
                        // counterpart): the population arriving here left this
                        // cell's mirror image, moving in the mirrored direction
                        int c[3] = {CV[q][0], CV[q][1], CV[q][2]};
                        if ((sx < 0 && (cfg.free_slip & AEROS_FACE_X0)) || (sx >= Nx && (cfg.free_slip & AEROS_FACE_X1))) { sx = x; c[0] = -c[0]; }
                        if ((sy < 0 && (cfg.free_slip & AEROS_FACE_Y0)) || (sy >= Ny && (cfg.free_slip & AEROS_FACE_Y1))) { sy = y; c[1] = -c[1]; }
                        if ((sz < 0 && (cfg.free_slip & AEROS_FACE_Z0)) || (sz >= Nz && (cfg.free_slip & AEROS_FACE_Z1))) { sz = z; c[2] = -c[2]; }
                        int qs = q;
                        for (int k = 0; k < Q; ++k)
                            if (CV[k][0] == c[0] && CV[k][1] == c[1] && CV[k][2] == c[2]) qs = k;
                        if (cfg.periodic_x) { if (sx < 0) sx += Nx; else if (sx >= Nx) sx -= Nx; }
                        if (cfg.periodic_z) { if (sz < 0) sz += Nz; else if (sz >= Nz) sz -= Nz; }
                        if (sx < 0 || sx >= Nx || sy < 0 || sy >= Ny || sz < 0 || sz >= Nz)
                            fNext[(size_t)d * Q + q] = f[(size_t)d * Q + q];
                        else {
                            const int s = idx(sx, sy, sz);
                            fNext[(size_t)d * Q + q] = (solid[s] || ground[s]) ? f[(size_t)d * Q + OPP[q]] : f[(size_t)s * Q + qs];
                        }
                    }
                }
            }
        if (!cfg.periodic_x) {   // inlet (aeros/src/lbm.cpp:1411-1463)
            const R u0 = inUx, u1 = inUy, u2 = inUz, us = u0 * u0 + u1 * u1 + u2 * u2;
            const int nA = ax == 0 ? Ny : Nx, nB = ax == 2 ? Ny : Nz;
            for (int b = 0; b < nB; ++b)
                for (int a = 0; a < nA; ++a) {
                    int cc;
                    if (ax == 0) cc = idx(sg > 0 ? 0 : Nx - 1, a, b);
                    else if (ax == 1) cc = idx(a, sg > 0 ? 0 : Ny - 1, b);
                    else cc = idx(a, b, sg > 0 ? 0 : Nz - 1);
                    if (solid[cc] || ground[cc]) continue;
                    for (int q = 0; q < Q; ++q) fNext[(size_t)cc * Q + q] = eq_guarded<R>(q, R(1), u0, u1, u2, us);
                }
        }
        if (cfg.convective_outlet && !cfg.periodic_x) {   // aeros/src/lbm.cpp:1465-1544
            float cc0 = U0mag * 0.8f;
            if (cc0 > 0.5f) cc0 = 0.5f;
            if (cc0 < 0.01f) cc0 = 0.08f;
            const R cf = (R)cc0, k2 = R(0.15) * (R)cc0;
            const int nA = ax == 0 ? Ny : Nx, nB = ax == 2 ? Ny : Nz;
            for (int b = 0; b < nB; ++b)
                for (int a = 0; a < nA; ++a) {
                    int co, ci, ci2;
                    if (ax == 0) { const int o = sg > 0 ? Nx - 1 : 0, i1 = sg > 0 ? Nx - 2 : 1, i2 = sg > 0 ? Nx - 3 : 2;
                                   co = idx(o, a, b); ci = idx(i1, a, b); ci2 = idx(i2, a, b); }
                    else if (ax == 1) { const int o = sg > 0 ? Ny - 1 : 0, i1 = sg > 0 ? Ny - 2 : 1, i2 = sg > 0 ? Ny - 3 : 2;
                                        co = idx(a, o, b); ci = idx(a, i1, b); ci2 = idx(a, i2, b); }
                    else { const int o = sg > 0 ? Nz - 1 : 0, i1 = sg > 0 ? Nz - 2 : 1, i2 = sg > 0 ? Nz - 3 : 2;
                           co = idx(a, b, o); ci = idx(a, b, i1); ci2 = idx(a, b, i2); }
                    if (co < 0 || co >= total || ci < 0 || ci >= total) continue;
                    if (solid[co] || ground[co]) continue;
                    for (int q = 0; q < Q; ++q) {
                        const R fO = f[(size_t)co * Q + q], fI = f[(size_t)ci * Q + q];
                        const R fI2 = (ci2 >= 0 && ci2 < total) ? f[(size_t)ci2 * Q + q] : fI;
                        fNext[(size_t)co * Q + q] = fO + cf * (fI - fO) + k2 * (fI - fI2);
                    }
                }
        }
    }

    // divergence reset (aeros/src/lbm.cpp:1583-1600)
    void reset_flow(float inUx, float inUy, float inUz) {
        resets++;
        const R u0 = inUx, u1 = inUy, u2 = inUz, us = u0 * u0 + u1 * u1 + u2 * u2;
        for (int cc = 0; cc < total; ++cc) {
            if (solid[cc] || ground[cc]) continue;
            for (int q = 0; q < Q; ++q) f[(size_t)cc * Q + q] = eq_guarded<R>(q, R(1), u0, u1, u2, us);
            quantize_cell(f, cc);
            rho[cc] = R(1); ux[cc] = u0; uy[cc] = u1; uz[cc] = u2;
        }
        avgRho = 1.0f;
        avgKE = (float)us;
    }

    // world velocity and pressure post-pass (aeros/src/lbm.cpp:1602-1705)
    void post_pass() {
        float a, b, c;
        inlet_velocity(cfg, a, b, c);
        float U0m = std::sqrt(a * a + b * b + c * c);
        if (!std::isfinite(U0m) || U0m < 1e-6f) U0m = 0.08f;
        float scale = cfg.flow_speed / U0m;
        if (!std::isfinite(scale) || scale > 1000.0f) scale = 20.0f;
        const R sc = (R)scale, rho0 = (R)cfg.air_density, cs2 = K<R>::cs2();
        long long bad = 0;
        for (int cc = 0; cc < total; ++cc) {
            if (solid[cc] || ground[cc]) { uxW[cc] = uyW[cc] = uzW[cc] = R(0); pres[cc] = R(0); continue; }
            R vx = ux[cc];
            if (!finite(vx)) vx = R(0);
            uxW[cc] = vx * sc; uyW[cc] = uy[cc] * sc; uzW[cc] = uz[cc] * sc;
            R r = rho[cc];
            if (!finite(r)) r = R(1);
            const R pl = cs2 * (r - R(1));
            R pp = pl * sc * sc * rho0;
            if (!finite(pp)) { bad++; pp = R(0); }
            pres[cc] = pp;
        }
        pBad = bad;
        int stride = 1;
        while (total / stride > 262144) stride *= 2;
        float mn = 1e30f, mx = -1e30f;
        const bool valid = ((int)prevRho.size() == total && prevTotal == total && prevStep >= 0 && curStep >= prevStep);
        if ((int)prevRho.size() != total) prevRho.assign((size_t)total, 1.0f);
        double ss = 0.0; int ns = 0;
        for (int cc = 0; cc < total; cc += stride) {
            const float p = (float)pres[cc];
            if (std::isfinite(p)) { if (p < mn) mn = p; if (p > mx) mx = p; }
            if (valid) {
                const float r = (float)rho[cc];
                if (std::isfinite(r)) { const float d = r - prevRho[cc]; ss += (double)d * (double)d; ++ns; }
            }
        }
        pMin = (mn <= mx && std::isfinite(mn)) ? mn : 0.0f;
        pMax = (mn <= mx && std::isfinite(mx)) ? mx : 0.0f;
        resRho = (valid && ns > 0) ? (float)std::sqrt(ss / (double)ns) : 0.0f;
        resU = std::fabs(maxVelLB - prevMaxVelLB);
        prevMaxVelLB = maxVelLB;
        for (int cc = 0; cc < total; ++cc) { const float r = (float)rho[cc]; prevRho[cc] = std::isfinite(r) ? r : 1.0f; }
        prevTotal = total; prevStep = curStep;
    }

    // lbmControlVolumeForce (aeros/src/lbm.cpp:2183-2232)
    void cv_force(float hxf, float hyf, float hzf, double out[3]) const {
        auto clampSide = [](int v, int n) { return v < 1 ? 1 : (v > n - 2 ? n - 2 : v); };
        const int cx = (int)(Nx * 0.5f), cy = (int)(Ny * 0.5f), cz = (int)(Nz * 0.5f);
        const int hx = std::max(2, (int)(Nx * hxf)), hy = std::max(2, (int)(Ny * hyf)), hz = std::max(2, (int)(Nz * hzf));
        const int x0 = clampSide(cx - hx, Nx), x1 = clampSide(cx + hx, Nx);
        const int y0 = clampSide(cy - hy, Ny), y1 = clampSide(cy + hy, Ny);
        const int z0 = clampSide(cz - hz, Nz), z1 = clampSide(cz + hz, Nz);
        double fx = 0, fy = 0, fz = 0;
        auto add = [&](int cc, R nx, R ny, R nz) {
            if (cc < 0 || cc >= total) return;
            if (solid[cc]) return;
            const R r = rho[cc], vx = ux[cc], vy = uy[cc], vz = uz[cc];
            const R p = (R(1) / R(3)) * (r - R(1));
            const R un = vx * nx + vy * ny + vz * nz;
            fx += (double)(p * nx + r * vx * un);
            fy += (double)(p * ny + r * vy * un);
            fz += (double)(p * nz + r * vz * un);
        };
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y) { add(idx(x0, y, z), R(-1), R(0), R(0)); add(idx(x1, y, z), R(1), R(0), R(0)); }
        for (int z = z0; z <= z1; ++z)
            for (int x = x0; x <= x1; ++x) { add(idx(x, y0, z), R(0), R(-1), R(0)); add(idx(x, y1, z), R(0), R(1), R(0)); }
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) { add(idx(x, y, z0), R(0), R(0), R(-1)); add(idx(x, y, z1), R(0), R(0), R(1)); }
        out[0] = -fx; out[1] = -fy; out[2] = -fz;
    }

    // lbmProjectedAreaCells (aeros/src/lbm.cpp:2236-2256)
    float projected_area(int axis) const {
        std::vector<char> hit((size_t)total, 0);
        double area = 0.0;
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y < Ny; ++y)
                for (int x = 0; x < Nx; ++x) {
                    if (!solid[idx(x, y, z)]) continue;
                    int a, b, aMax;
                    if (axis == 0) { a = y; b = z; aMax = Ny; }
                    else if (axis == 1) { a = x; b = z; aMax = Nx; }
                    else { a = x; b = y; aMax = Nx; }
                    const size_t k = (size_t)b * (size_t)aMax + (size_t)a;
                    if (k < hit.size() && !hit[k]) { hit[k] = 1; area += 1.0; }
                }
        return (float)area;
    }

    // bench checksum (aeros/tests/bench_lbm.cpp:103-107): sequential double sums
    void checksums(double* sr, double* su) const {
        double a = 0, b = 0;
        for (int cc = 0; cc < total; ++cc) { a += (double)rho[cc]; b += (double)(ux[cc] + uy[cc] + uz[cc]); }
        *sr = a; *su = b;
    }

    const std::vector<R>* field(int id) const {
        switch (id) {
            case AEROS_FIELD_RHO: return &rho; case AEROS_FIELD_UX: return &ux; case AEROS_FIELD_UY: return &uy;
            case AEROS_FIELD_UZ: return &uz; case AEROS_FIELD_UX_WORLD: return &uxW; case AEROS_FIELD_UY_WORLD: return &uyW;
            case AEROS_FIELD_UZ_WORLD: return &uzW; case AEROS_FIELD_PRESSURE: return &pres; case AEROS_FIELD_TKE: return &tke;
            case AEROS_FIELD_STRAIN: return &strain; case AEROS_FIELD_T: return thermalReady ? &T : nullptr;
            case AEROS_FIELD_BUOYANCY: return thermalReady ? &buoy : nullptr;
            default: return nullptr;
        }
    }
};

struct Handle {
    int prec = AEROS_FP32;
    std::unique_ptr<Sim<float>> s32;
    std::unique_ptr<Sim<double>> s64;
};

template <class S> void fill_scalars(const S& s, aeros_lbm_scalars* o) {
    std::memset(o, 0, sizeof(*o));
    o->step = s.curStep; o->converged = s.converged; o->resets = s.resets;
    o->nx = s.Nx; o->ny = s.Ny; o->nz = s.Nz;
    o->avg_rho = s.avgRho; o->avg_ke = s.avgKE; o->max_u_lb = s.maxVelLB; o->max_u_world = s.maxVelW;
    o->tke = s.tkeMean; o->convergence = s.conv;
    o->neg_pops = s.negPops; o->bad_pops = s.badPops; o->rho_out_of_range = s.rhoOut;
    o->rho_non_finite = s.rhoNonFin; o->u_limited = s.uLim; o->pressure_non_finite = s.pBad;
    o->tau = s.tau; o->viscosity = s.viscosity; o->U0 = s.U0; o->reynolds_phys = s.rePhys; o->reynolds_lb = s.reLB;
    o->blockage = s.blockage; o->pressure_min = s.pMin; o->pressure_max = s.pMax;
    o->residual_rho = s.resRho; o->residual_u = s.resU; o->fluid_cells = s.fluidCells;
    o->nusselt = s.tNusselt; o->thermal_mean_t = s.tMeanT; o->thermal_alpha = s.tAlpha; o->thermal_gravity = s.tG;
    o->thermal_tau = s.tauT; o->thermal_rayleigh = s.tRa; o->thermal_prandtl = s.tPr; o->thermal_ready = s.thermalReady;
    o->thermal_ylo = s.yLo; o->thermal_yhi = s.yHi;
}

}  // namespace

// ============================================================================ C API
extern "C" {

void ora_config_defaults(aeros_lbm_config* c) {
    // LBMParams (aeros/src/lbm.h:11-75), ThermalParams (aeros/src/thermal_lbm.h:33-41),
    // globals (aeros/src/globals.cpp:59-145, aeros/src/atmosphere.cpp:9-10)
    std::memset(c, 0, sizeof(*c));
    c->struct_size = (int32_t)sizeof(*c);
    c->voxel_resolution = 48;
    c->vox_min[0] = c->vox_min[1] = c->vox_min[2] = -1.0f;
    c->vox_max[0] = c->vox_max[1] = c->vox_max[2] = 1.0f;
    c->flow_speed = 2.0f; c->air_density = 1.225f; c->air_temperature = 288.15f; c->max_dim = 1.0f;
    c->tau = 0.6f; c->U0 = 0.1f; c->smagorinsky_c = 0.12f; c->convergence_threshold = 1e-5f;
    c->domain_padding = 0.25f; c->inlet_turbulence = 0.02f;
    c->use_les = 1; c->use_trt = 0; c->convective_outlet = 1; c->auto_viscosity = 1;
    c->rayleigh = 1.0e5f; c->prandtl = 0.71f; c->fixed_wall_temps = 1; c->hot_temp = 1.0f; c->cold_temp = 0.0f;
    c->perturbation = 0.02f;
    c->precision = AEROS_FP32;
}

void* ora_create(const aeros_lbm_config* c, int* err) {
    auto h = new (std::nothrow) Handle();
    if (!h) { if (err) *err = AEROS_ERR_NOMEM; return nullptr; }
    h->prec = c->precision;
    int rc;
    if (c->precision == AEROS_FP64 || c->precision == AEROS_FP32_FP64C) {
        h->s64.reset(new Sim<double>());
        h->s64->storage = c->precision;
        rc = h->s64->init(*c);
    }
    else { h->s32.reset(new Sim<float>()); h->s32->storage = c->precision; rc = h->s32->init(*c); }
    if (err) *err = rc;
    if (rc != AEROS_OK) { delete h; return nullptr; }
    return h;
}

void ora_destroy(void* p) { delete (Handle*)p; }

#define ORA_DISPATCH(h, expr) ((h)->s64 ? [&](auto& S) { return expr; }(*(h)->s64) : [&](auto& S) { return expr; }(*(h)->s32))

int ora_dims(void* p, int32_t d[3], float mn[3], float cs[3]) {
    Handle* h = (Handle*)p;
    return ORA_DISPATCH(h, (d[0] = S.Nx, d[1] = S.Ny, d[2] = S.Nz, mn[0] = S.minX, mn[1] = S.minY, mn[2] = S.minZ,
                            cs[0] = S.csX, cs[1] = S.csY, cs[2] = S.csZ, 0));
}

int ora_step(void* p, int n) { Handle* h = (Handle*)p; return ORA_DISPATCH(h, (S.step(n), 0)); }

int ora_set_mask(void* p, const uint8_t* m) { Handle* h = (Handle*)p; return ORA_DISPATCH(h, S.set_mask(m)); }

int ora_read_field(void* p, int id, double* out) {
    Handle* h = (Handle*)p;
    return ORA_DISPATCH(h, [&]() {
        if (id == AEROS_FIELD_SOLID || id == AEROS_FIELD_GROUND || id == AEROS_FIELD_WALLDIST) {
            for (int c = 0; c < S.total; ++c)
                out[c] = id == AEROS_FIELD_SOLID ? S.solid[c] : (id == AEROS_FIELD_GROUND ? S.ground[c] : S.wallDist[c]);
            return 0;
        }
        auto* v = S.field(id);
        if (!v) return AEROS_ERR_ARG;
        for (int c = 0; c < S.total; ++c) out[c] = (double)(*v)[c];
        return 0;
    }());
}

int ora_read_populations(void* p, double* out) {
    Handle* h = (Handle*)p;
    return ORA_DISPATCH(h, [&]() { for (size_t i = 0; i < S.f.size(); ++i) out[i] = (double)S.f[i]; return 0; }());
}

int ora_scalars(void* p, aeros_lbm_scalars* o) { Handle* h = (Handle*)p; return ORA_DISPATCH(h, (fill_scalars(S, o), 0)); }

int ora_cv_force(void* p, float a, float b, float c, double out[3]) {
    Handle* h = (Handle*)p; return ORA_DISPATCH(h, (S.cv_force(a, b, c, out), 0));
}

float ora_projected_area(void* p, int axis) { Handle* h = (Handle*)p; return ORA_DISPATCH(h, S.projected_area(axis)); }

int ora_checksums(void* p, double* r, double* u) { Handle* h = (Handle*)p; return ORA_DISPATCH(h, (S.checksums(r, u), 0)); }

int ora_thermal_upload(void* p, const float* t) {
    Handle* h = (Handle*)p; return ORA_DISPATCH(h, S.thermal_upload(t) ? 0 : AEROS_ERR_STATE);
}

int ora_thermal_reset(void* p) { Handle* h = (Handle*)p; return ORA_DISPATCH(h, (S.thermal_reset(), 0)); }

// lbmEquilibriumMoments (aeros/src/lbm.cpp:249-266)
int ora_equilibrium_moments(float rho, float ux, float uy, float uz, float* s, float* mx, float* my, float* mz) {
    if (!s || !mx || !my || !mz) return 0;
    if (!std::isfinite(rho) || !std::isfinite(ux) || !std::isfinite(uy) || !std::isfinite(uz)) return 0;
    const float us = ux * ux + uy * uy + uz * uz;
    float m = 0, px = 0, py = 0, pz = 0;
    for (int q = 0; q < Q; ++q) {
        const float fe = eq_guarded<float>(q, rho, ux, uy, uz, us);
        if (!std::isfinite(fe)) return 0;
        m += fe; px += fe * CV[q][0]; py += fe * CV[q][1]; pz += fe * CV[q][2];
    }
    *s = m; *mx = px; *my = py; *mz = pz;
    return 1;
}

// avgRho and step of the last divergence reset (diagnostic of the reset path)
float ora_last_divergence(void* p, int* step) {
    Handle* h = (Handle*)p;
    return ORA_DISPATCH(h, (*step = S.lastDivergedStep, S.lastDivergedRho));
}

float ora_sutherland(float t_kelvin) { return sutherland(t_kelvin); }
float ora_q_half(float x) { return q_half(x); }
// The values the stored populations are taken relative to (w_q, or the inflow
// equilibrium for the equilibrium-shifted storages).
int ora_storage_shift(void* p, double* out) {
    Handle* h = (Handle*)p;
    const bool eq = h->prec == AEROS_FP16_EQ || h->prec == AEROS_BF16_EQ;
    for (int q = 0; q < Q; ++q)
        out[q] = eq ? (double)h->s32->eqs[q] : (h->s64 ? (double)K<double>::w(q) : (double)K<float>::w(q));
    return 0;
}
float ora_q_bf16(float x) { return q_bf16(x); }

// Voxelize on an arbitrary grid (aeros/src/voxel_grid.cpp:313-351 restated):
// drop non-finite and degenerate triangles, build the BVH, vote at centres.
int ora_voxelize(const float* v, int64_t ntris, int32_t nx, int32_t ny, int32_t nz, const float mn[3],
                 const float cs[3], int32_t* out) {
    Bvh b;
    b.tris.reserve((size_t)ntris);
    for (int64_t t = 0; t < ntris; ++t) {
        const float* p = v + t * 9;
        bool ok = true;
        for (int k = 0; k < 9; ++k) ok = ok && std::isfinite(p[k]);
        if (!ok) continue;
        Tri tr;
        tr.v0 = {p[0], p[1], p[2]}; tr.v1 = {p[3], p[4], p[5]}; tr.v2 = {p[6], p[7], p[8]};
        tr.lo = vmin(tr.v0, vmin(tr.v1, tr.v2)); tr.hi = vmax(tr.v0, vmax(tr.v1, tr.v2));
        tr.cen = {(tr.v0.x + tr.v1.x + tr.v2.x) * (1.0f / 3.0f), (tr.v0.y + tr.v1.y + tr.v2.y) * (1.0f / 3.0f),
                  (tr.v0.z + tr.v1.z + tr.v2.z) * (1.0f / 3.0f)};
        // isDegenerate (aeros/src/mesh_bvh.h:51-54): |cross(v1-v0, v2-v0)| <= 1e-14
        const float ax = tr.v1.x - tr.v0.x, ay = tr.v1.y - tr.v0.y, az = tr.v1.z - tr.v0.z;
        const float bx = tr.v2.x - tr.v0.x, by = tr.v2.y - tr.v0.y, bz = tr.v2.z - tr.v0.z;
        const float cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
        const float a2 = std::sqrt(cx * cx + cy * cy + cz * cz);
        if (!(a2 > 1e-14f) || !std::isfinite(a2)) continue;
        b.tris.push_back(tr);
    }
    if (!b.tris.empty()) b.build(0, (int)b.tris.size(), 0);
    const V3 cell{cs[0], cs[1], cs[2]};
    #pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const V3 p{mn[0] + (i + 0.5f) * cs[0], mn[1] + (j + 0.5f) * cs[1], mn[2] + (k + 0.5f) * cs[2]};
                out[((size_t)k * ny + j) * nx + i] = (!b.tris.empty() && b.inside(p, cell)) ? 1 : 0;
            }
    return 0;
}

int ora_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n; return 1;
#endif
}

}  // extern "C"
