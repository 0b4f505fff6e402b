// Host half of initLBM (aeros/src/lbm.cpp:405-886): domain from the voxel box
// plus padding, mask transfer by cell-centre world coordinates, ground slab,
// u8 wall distance, tau from Re or from the user, thermal parameters.
// Everything here runs once per case; the per-step work is on the GPU.
//
// Float expressions keep the reference's operand order so that the domain,
// tau and the inlet velocity are bit-identical to the reference's.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "lbm_common.h"

namespace aeros {

namespace {

// glm::radians<float>
inline float deg2rad(float d) { return d * 0.01745329251994329576923690768489f; }

struct Dir { float d[3]; };

// unit flow direction from azimuth/elevation (aeros/src/lbm.cpp:201-214)
Dir flow_direction(const aeros_lbm_config& c) {
    const float az = deg2rad(c.flow_azimuth_deg), el = deg2rad(c.flow_elevation_deg);
    const float ce = cosf(el), se = sinf(el), ca = cosf(az), sa = sinf(az);
    Dir r{{ce * ca, se, ce * sa}};
    const float len = std::sqrt(r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2]);
    if (len < 1e-6f || !std::isfinite(len)) { r.d[0] = 1; r.d[1] = 0; r.d[2] = 0; }
    else for (float& v : r.d) v /= len;
    return r;
}

// airSutherlandViscosity (aeros/src/flow_field.h:46-54)
float sutherland_mu(float Tk) {
    if (!std::isfinite(Tk) || Tk < 50.0f) Tk = 288.15f;
    Tk = std::min(Tk, 2000.0f);
    const float T0 = 273.15f, S = 110.4f, mu0 = 1.716e-5f;
    float mu = mu0 * std::pow(Tk / T0, 1.5f) * (T0 + S) / (Tk + S);
    return (std::isfinite(mu) && mu >= 1e-7f) ? mu : 1.81e-5f;
}

void set_msg(char* msg, int len, const char* text) {
    if (msg && len > 0) std::snprintf(msg, (size_t)len, "%s", text);
}

}  // namespace

void config_defaults(aeros_lbm_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->struct_size = (int32_t)sizeof(*c);
    c->voxel_resolution = 48;                       // voxelResolution (aeros/src/globals.cpp:119)
    for (int i = 0; i < 3; ++i) { c->vox_min[i] = -1.0f; c->vox_max[i] = 1.0f; }
    c->flow_speed = 2.0f;                           // aeros/src/globals.cpp:59
    c->air_density = 1.225f;                        // aeros/src/atmosphere.cpp:9
    c->air_temperature = 288.15f;                   // aeros/src/atmosphere.cpp:10
    c->max_dim = 1.0f;                              // aeros/src/globals.cpp:107
    c->tau = 0.6f; c->U0 = 0.1f; c->smagorinsky_c = 0.12f;   // LBMParams (aeros/src/lbm.h:11-75)
    c->convergence_threshold = 1e-5f; c->domain_padding = 0.25f; c->inlet_turbulence = 0.02f;
    c->use_les = 1; c->use_trt = 0; c->convective_outlet = 1; c->auto_viscosity = 1;
    c->rayleigh = 1.0e5f; c->prandtl = 0.71f; c->fixed_wall_temps = 1;   // ThermalParams
    c->hot_temp = 1.0f; c->cold_temp = 0.0f; c->perturbation = 0.02f;
    c->precision = AEROS_FP32;
}

// computeInletVelocityLB (aeros/src/lbm.cpp:186-199)
void inlet_velocity(const aeros_lbm_config& c, float& ux, float& uy, float& uz) {
    const float az = deg2rad(c.flow_azimuth_deg), el = deg2rad(c.flow_elevation_deg);
    const float ce = cosf(el), se = sinf(el), ca = cosf(az), sa = sinf(az);
    float d[3] = {ce * ca, se, ce * sa};
    float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len < 1e-6f || !std::isfinite(len)) { d[0] = 1; d[1] = 0; d[2] = 0; len = 1; }
    for (float& v : d) v /= len;
    float U = c.U0;
    if (!std::isfinite(U) || U < 1e-6f) U = 0.08f;
    if (U > 0.15f) U = 0.15f;
    ux = d[0] * U; uy = d[1] * U; uz = d[2] * U;
}

// flowParams.vx/vy/vz, the free stream in world units (aeros/src/flow_field.cpp:49-68)
void free_stream(const aeros_lbm_config& c, float v[3]) {
    float speed = std::isfinite(c.flow_speed) ? c.flow_speed : 2.0f;
    if (speed < 0) speed = 0;
    if (speed > 1000.0f) speed = 1000.0f;
    const float az0 = std::isfinite(c.flow_azimuth_deg) ? c.flow_azimuth_deg : 0.0f;
    float el0 = std::isfinite(c.flow_elevation_deg) ? c.flow_elevation_deg : 0.0f;
    if (el0 < -89.0f) el0 = -89.0f;
    if (el0 > 89.0f) el0 = 89.0f;
    const float az = deg2rad(az0), el = deg2rad(el0);
    const float ce = cosf(el), se = sinf(el), ca = cosf(az), sa = sinf(az);
    float d[3] = {ce * ca, se, ce * sa};
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len < 1e-6f || !std::isfinite(len)) { d[0] = 1; d[1] = 0; d[2] = 0; }
    else for (float& x : d) x /= len;
    for (int k = 0; k < 3; ++k) v[k] = d[k] * speed;
}

void flow_axis(const aeros_lbm_config& c, int& axis, int& sign) {
    const Dir r = flow_direction(c);
    const float a0 = std::fabs(r.d[0]), a1 = std::fabs(r.d[1]), a2 = std::fabs(r.d[2]);
    if (a0 >= a1 && a0 >= a2) axis = 0;
    else if (a1 >= a2) axis = 1;
    else axis = 2;
    sign = r.d[axis] > 0 ? 1 : -1;
}

int check_boundaries(const aeros_lbm_config& c, char* msg, int msglen) {
    if ((c.free_slip & ~63) || (c.periodic_x && (c.free_slip & (AEROS_FACE_X0 | AEROS_FACE_X1))) ||
        (c.periodic_z && (c.free_slip & (AEROS_FACE_Z0 | AEROS_FACE_Z1)))) {
        set_msg(msg, msglen, "free_slip names an unknown face or a face of a periodic axis");
        return AEROS_ERR_ARG;
    }
    // with the thermal lattice, free-slip walls act on the D3Q19 lattice only:
    // the D3Q7 lattice keeps its own shell rules (x and z wrap, y walls)
    return AEROS_OK;
}

int setup_domain(const aeros_lbm_config& cfg, Setup& s, char* msg, int msglen) {
    s = Setup{};
    s.cfg = cfg;
    const aeros_lbm_config& c = cfg;
    if (int rc = check_boundaries(c, msg, msglen)) return rc;
    const bool vox = c.voxels && c.vox_n[0] > 0 && c.vox_n[1] > 0 && c.vox_n[2] > 0;
    const int64_t cap_pad = c.cell_cap > 0 ? c.cell_cap : 8LL * 1024 * 1024;    // aeros/src/lbm.cpp:467
    const int64_t cap_hard = c.cell_cap > 0 ? c.cell_cap : 20LL * 1024 * 1024;  // aeros/src/lbm.cpp:558
    int n[3];
    if (!vox) {
        n[0] = n[1] = n[2] = c.voxel_resolution;
        s.min[0] = s.min[1] = s.min[2] = -1.0f;
    } else {
        // padded external domain: 2L upstream, 4L downstream, 2.5L lateral (aeros/src/lbm.cpp:430-514)
        float cs[3];
        for (int i = 0; i < 3; ++i) {
            cs[i] = (c.vox_max[i] - c.vox_min[i]) / std::max(1, c.vox_n[i]);
            if (!std::isfinite(cs[i]) || cs[i] < 1e-6f) cs[i] = 0.1f;
        }
        float L = std::max(std::max(c.vox_max[0] - c.vox_min[0], c.vox_max[1] - c.vox_min[1]),
                           c.vox_max[2] - c.vox_min[2]);
        if (!std::isfinite(L) || L < 1e-6f) L = 1.0f;
        const Dir dir = flow_direction(c);
        float pad = std::isfinite(c.domain_padding) ? c.domain_padding : 1.0f;
        pad = std::min(std::max(pad, 0.0f), 1.0f);
        float lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
            const float d = std::isfinite(dir.d[i]) ? dir.d[i] : 0.0f;
            const float ad = std::fabs(d);
            const float up = 2.0f * L * ad * pad, down = 4.0f * L * ad * pad;
            const float lat = 2.5f * L * (1.0f - ad) * pad;
            lo[i] = std::max(d > 0.0f ? up : down, lat);
            hi[i] = std::max(d > 0.0f ? down : up, lat);
        }
        auto cells_at = [&](float k) {
            int64_t t = 1;
            for (int i = 0; i < 3; ++i) {
                const int extra = std::max(0, (int)std::lround((lo[i] + hi[i]) * k / cs[i]));
                t *= (int64_t)(c.vox_n[i] + extra);
            }
            return t;
        };
        float k = 1.0f;
        while (k > 0.05f && cells_at(k) > cap_pad) k *= 0.75f;
        if (cells_at(k) > cap_pad) k = 0.0f;
        for (int i = 0; i < 3; ++i) {
            s.min[i] = c.vox_min[i] - lo[i] * k;
            int eLo = (int)std::ceil((lo[i] * k - 1e-9f) / cs[i]);
            int eHi = (int)std::ceil((hi[i] * k - 1e-9f) / cs[i]);
            if (k <= 0.0f) eLo = eHi = 0;
            n[i] = std::max(8, c.vox_n[i] + std::max(0, eLo) + std::max(0, eHi));
            s.cell[i] = cs[i];
        }
    }
    s.nx = n[0]; s.ny = n[1]; s.nz = n[2];
    const int64_t tot = (int64_t)n[0] * n[1] * n[2];
    if (tot > cap_pad) {   // the reference refuses instead of shrinking (aeros/src/lbm.cpp:526-545)
        char b[256];
        std::snprintf(b, sizeof b, "domain %dx%dx%d = %lld cells exceeds the cap %lld", n[0], n[1], n[2],
                      (long long)tot, (long long)cap_pad);
        set_msg(msg, msglen, b);
        return AEROS_ERR_TOO_LARGE;
    }
    for (int i = 0; i < 3; ++i)
        if (!(s.cell[i] > 1e-6f) || !std::isfinite(s.cell[i])) s.cell[i] = 2.0f / std::max(1, n[i]);
    if (tot <= 0 || tot > cap_hard) { set_msg(msg, msglen, "invalid total cell count"); return AEROS_ERR_TOO_LARGE; }
    s.total = tot;
    try {
        s.solid.assign((size_t)tot, 0);
        s.ground.assign((size_t)tot, 0);
    } catch (...) { set_msg(msg, msglen, "host allocation failed"); return AEROS_ERR_NOMEM; }

    if (vox) {   // mask by cell-centre world coordinates (aeros/src/lbm.cpp:596-622)
        float b[3];
        for (int i = 0; i < 3; ++i) {
            b[i] = (c.vox_max[i] - c.vox_min[i]) / (float)c.vox_n[i];
            if (!(b[i] > 1e-9f)) b[i] = 1e-9f;
        }
        std::vector<int> gx((size_t)s.nx), gy((size_t)s.ny), gz((size_t)s.nz);
        int* g[3] = {gx.data(), gy.data(), gz.data()};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < n[i]; ++j) {
                const float p = s.min[i] + (j + 0.5f) * s.cell[i];
                g[i][j] = (int)std::floor((p - c.vox_min[i]) / b[i]);
            }
        for (int z = 0; z < s.nz; ++z) {
            if (gz[z] < 0 || gz[z] >= c.vox_n[2]) continue;
            for (int y = 0; y < s.ny; ++y) {
                if (gy[y] < 0 || gy[y] >= c.vox_n[1]) continue;
                const int32_t* vrow = c.voxels + ((int64_t)gz[z] * c.vox_n[1] + gy[y]) * c.vox_n[0];
                uint8_t* srow = s.solid.data() + ((int64_t)z * s.ny + y) * s.nx;
                for (int x = 0; x < s.nx; ++x)
                    if (gx[x] >= 0 && gx[x] < c.vox_n[0]) srow[x] = vrow[gx[x]] == 1 ? 1 : 0;
            }
        }
    }
    apply_ground(s);
    return AEROS_OK;
}

void apply_ground(Setup& s) {
    const aeros_lbm_config& c = s.cfg;
    if (!(c.ground_effect || c.use_ground)) return;   // ground slab (aeros/src/lbm.cpp:624-643)
    float gyw = c.vox_min[1] + c.ground_effect_height;
    if (c.use_ground) gyw = c.vox_min[1] + c.ground_height;
    if (!std::isfinite(gyw)) gyw = c.vox_min[1];
    int gy0 = (int)((gyw - s.min[1]) / s.cell[1]);
    gy0 = std::min(std::max(gy0, 0), s.ny - 1);
    for (int z = 0; z < s.nz; ++z)
        for (int y = 0; y <= gy0; ++y)
            for (int x = 0; x < s.nx; ++x) {
                const int64_t cc = ((int64_t)z * s.ny + y) * s.nx + x;
                if (!s.solid[cc]) s.ground[cc] = 1;
            }
}

int setup_after_mask(Setup& s, char* msg, int msglen) {
    const aeros_lbm_config& c = s.cfg;
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const int64_t tot = s.total, sy = nx, sz = (int64_t)nx * ny;
    // u8 wall distance: forward then backward 6-neighbour chamfer, saturating
    // at 255 (aeros/src/lbm.cpp:647-675)
    s.wall.assign((size_t)tot, 255);
    for (int64_t i = 0; i < tot; ++i) if (s.solid[i] || s.ground[i]) s.wall[i] = 0;
    uint8_t* w = s.wall.data();
    auto relax = [](uint8_t& d, uint8_t nb) { if ((unsigned)nb + 1u < d) d = (uint8_t)(nb + 1); };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            uint8_t* r = w + (z * sz + (int64_t)y * sy);
            for (int x = 0; x < nx; ++x) {
                if (r[x] == 0) continue;
                uint8_t d = r[x];
                if (x > 0) relax(d, r[x - 1]);
                if (y > 0) relax(d, r[x - sy]);
                if (z > 0) relax(d, r[x - sz]);
                r[x] = d;
            }
        }
    for (int z = nz - 1; z >= 0; --z)
        for (int y = ny - 1; y >= 0; --y) {
            uint8_t* r = w + (z * sz + (int64_t)y * sy);
            for (int x = nx - 1; x >= 0; --x) {
                if (r[x] == 0) continue;
                uint8_t d = r[x];
                if (x < nx - 1) relax(d, r[x + 1]);
                if (y < ny - 1) relax(d, r[x + sy]);
                if (z < nz - 1) relax(d, r[x + sz]);
                r[x] = d;
            }
        }
    int64_t fluid = 0;
    for (int64_t i = 0; i < tot; ++i) fluid += (!s.solid[i] && !s.ground[i]);
    s.fluid = std::max<int64_t>(fluid, 1);

    inlet_velocity(c, s.inlet[0], s.inlet[1], s.inlet[2]);
    flow_axis(c, s.axis, s.sign);
    s.curl_noise = c.inlet_turbulence > 0.001f;

    // tau from the physical Re or from the user (aeros/src/lbm.cpp:741-815)
    const float mu = sutherland_mu(c.air_temperature);
    const float L = c.max_dim < 1e-6f ? 1.0f : c.max_dim;
    const float rho_air = (!std::isfinite(c.air_density) || c.air_density < 0.01f) ? 1.225f : c.air_density;
    float U0 = std::sqrt(s.inlet[0] * s.inlet[0] + s.inlet[1] * s.inlet[1] + s.inlet[2] * s.inlet[2]);
    if (!std::isfinite(U0) || U0 < 1e-6f) U0 = 0.08f;
    float re_phys = rho_air * c.flow_speed * L / mu;
    if (!std::isfinite(re_phys) || re_phys < 1.0f) re_phys = 1e4f;
    float cs_avg = (s.cell[0] + s.cell[1] + s.cell[2]) / 3.0f;
    if (cs_avg < 1e-6f) cs_avg = L / nx;
    const float L_lb = std::min(std::max(L / cs_avg, 5.0f), 200.0f);
    const float cs2 = 1.0f / 3.0f;
    float tau_phys = ((U0 * L_lb) / re_phys) / cs2 + 0.5f;
    if (!std::isfinite(tau_phys)) tau_phys = 1.5f;
    float tau;
    if (c.auto_viscosity) tau = std::min(std::max(tau_phys, 0.51f), 1.5f);
    else {
        tau = std::isfinite(c.tau) ? c.tau : 0.6f;
        tau = std::min(std::max(tau, 0.51f), 1.5f);
    }
    s.tau = tau;
    s.viscosity = (tau - 0.5f) * cs2;
    s.U0 = U0;
    s.cfg.tau = tau;   // initLBM writes lbmParams.tau and lbmParams.U0 back
    s.cfg.U0 = U0;
    const float nu = (tau - 0.5f) * cs2;
    s.re_lb = nu > 1e-12f ? U0 * L_lb / nu : 0.0f;
    if (!std::isfinite(s.re_lb)) s.re_lb = 0.0f;
    s.re_phys = re_phys;
    s.blockage = 0.0f;
    if (c.frontal_area > 0.0f) {   // aeros/src/lbm.cpp:826-853
        const float sec = s.axis == 0 ? (ny * s.cell[1]) * (nz * s.cell[2])
                        : (s.axis == 1 ? (nx * s.cell[0]) * (nz * s.cell[2]) : (nx * s.cell[0]) * (ny * s.cell[1]));
        if (std::isfinite(sec) && sec > 1e-9f && std::isfinite(c.frontal_area)) s.blockage = c.frontal_area / sec;
    }

    // thermalInit (aeros/src/thermal_lbm.cpp:84-140)
    s.thermal_ready = false;
    if (c.thermal) {
        bool ok = nx >= 4 && ny >= 4 && nz >= 1 && (c.cell_cap > 0 || tot <= 20LL * 1024 * 1024);
        float nuT = s.viscosity;
        if (!std::isfinite(nuT) || nuT <= 0.0f) nuT = 0.05f;
        s.ylo = -1; s.yhi = -1;
        if (ok) {
            for (int y = 0; y < ny; ++y) {
                bool any = false;
                for (int z = 0; z < nz && !any; ++z) {
                    const int64_t base = ((int64_t)z * ny + y) * nx;
                    for (int x = 0; x < nx; ++x)
                        if (!s.solid[base + x] && !s.ground[base + x]) { any = true; break; }
                }
                if (any) { if (s.ylo < 0) s.ylo = y; s.yhi = y; }
            }
            ok = s.ylo >= 0 && s.yhi > s.ylo;
        }
        if (ok) {
            s.H = (float)(s.yhi - s.ylo + 1);
            s.prandtl = std::max(0.05f, c.prandtl);
            s.alpha = nuT / s.prandtl;
            s.tauT = 0.5f + s.alpha / 0.25f;
            if (!std::isfinite(s.tauT) || s.tauT < 0.505f) s.tauT = 0.505f;
            if (s.tauT > 1.5f) { s.tauT = 1.5f; s.alpha = 0.25f * (s.tauT - 0.5f); s.prandtl = nuT / s.alpha; }
            const float dT = c.hot_temp - c.cold_temp;
            s.gravity = c.rayleigh * nuT * s.alpha / (std::max(1e-6f, std::fabs(dT)) * s.H * s.H * s.H);
            s.rayleigh = c.rayleigh;   // computeDiagnostics reports the configured values
            s.prandtl = c.prandtl;
            s.thermal_ready = true;
        } else {
            s.cfg.thermal = 0;   // initLBM turns the module off (aeros/src/lbm.cpp:880-884)
        }
    }
    (void)msg; (void)msglen;
    return AEROS_OK;
}

// thermalResetField's temperature (aeros/src/thermal_lbm.cpp:157-180)
void thermal_initial_field(const Setup& s, std::vector<float>& T) {
    const aeros_lbm_config& c = s.cfg;
    const float Th = c.hot_temp, Tc = c.cold_temp, amp = c.perturbation * (Th - Tc);
    T.assign((size_t)s.total, 0.0f);
    for (int z = 0; z < s.nz; ++z)
        for (int y = 0; y < s.ny; ++y)
            for (int x = 0; x < s.nx; ++x) {
                const int64_t cc = ((int64_t)z * s.ny + y) * s.nx + x;
                float t;
                if (s.solid[cc] || s.ground[cc]) {
                    if (c.fixed_wall_temps && y <= 0) t = Th;
                    else if (c.fixed_wall_temps && y >= s.ny - 1) t = Tc;
                    else t = 0.5f * (Th + Tc);
                } else {
                    const float sl = (float)(y - s.ylo) / std::max(1.0f, s.H - 1.0f);
                    t = Th + (Tc - Th) * sl;
                    const float px = 6.2831853f * (float)x / (float)s.nx;
                    const float py = 3.1415926f * sl;
                    t += amp * std::cos(px) * std::sin(py);
                }
                T[cc] = t;
            }
}

// curl-noise inflow of initLBM (aeros/src/lbm.cpp:698-726)
void curl_noise_velocity(const Setup& s, std::vector<float>& u3) {
    const float in0 = s.inlet[0], in1 = s.inlet[1], in2 = s.inlet[2];
    const float turb = s.cfg.inlet_turbulence * 0.04f;
    u3.assign((size_t)s.total * 3, 0.0f);
    const float flen = sqrtf(in0 * in0 + in1 * in1 + in2 * in2);
    for (int64_t cc = 0; cc < s.total; ++cc) {
        if (s.solid[cc] || s.ground[cc]) continue;
        float u[3] = {in0, in1, in2};
        const int x = (int)(cc % s.nx), y = (int)((cc / s.nx) % s.ny), z = (int)(cc / ((int64_t)s.nx * s.ny));
        const float fx = (float)x / s.nx * 6.28f, fy = (float)y / s.ny * 6.28f, fz = (float)z / s.nz * 6.28f;
        const float a = cosf(fx * 1.2f + fy * 0.8f), b = cosf(fz * 1.1f + fx * 0.9f), d = cosf(fy * 1.3f + fz * 0.7f);
        const float r0 = (a * 0.8f - b * 1.1f) * turb, r1 = (d * 0.7f - a * 1.2f) * turb, r2 = (b * 0.9f - d * 1.3f) * turb;
        if (flen > 1e-6f) {
            const float e0 = in0 / flen, e1 = in1 / flen, e2 = in2 / flen;
            const float along = r0 * e0 + r1 * e1 + r2 * e2;
            u[0] += r0 - e0 * along; u[1] += r1 - e1 * along; u[2] += r2 - e2 * along;
        }
        u3[cc * 3 + 0] = u[0]; u3[cc * 3 + 1] = u[1]; u3[cc * 3 + 2] = u[2];
    }
}

}  // namespace aeros
