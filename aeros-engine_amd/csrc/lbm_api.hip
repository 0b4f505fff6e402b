// C ABI of the MI355X LBM solver (include/aeros_lbm.h): handle lifetime,
// device memory, the per-call step loop with the reference's divergence
// handling, readback, forces, thermal upload and diagnostics.
//
// The step loop keeps the reference's semantics (aeros/src/lbm.cpp:907-1709):
// populations are stored after collision, so the kernel of reference step s
// pulls from the state written by step s-1; the first step after an
// initialisation or a reset collides in place.  All per-step decisions (NaN
// storm, divergence, convergence) are taken on the device by the last block
// of the step kernel; the host reads one small status record per call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "lbm_kernels.hpp"

using namespace aeros;

// ------------------------------------------------------------ errors
namespace {
thread_local int t_err_code = 0;
thread_local char t_err_msg[1024] = {0};

int set_err(int code, const char* fmt, ...) {
    t_err_code = code;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(t_err_msg, sizeof t_err_msg, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return set_err(AEROS_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return set_err(AEROS_ERR_NOMEM, "hipMalloc of %zu bytes failed", b);
        }
        bytes = b;
        return AEROS_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

inline int grid_for(int64_t n, int block = 256) {
    int64_t g = (n + block - 1) / block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

// Step-kernel blocks: each owns a run of consecutive 256-cell rows.  Short
// runs keep the blocks that are resident at one time on a compact window of
// the lattice and make the end of the launch even; longer runs amortise the
// block's reduction.  Measured best: 4 rows from 8M cells up (33.5M cells:
// 4 rows 1.19 ms, 2 rows 1.21, 8 rows 1.25), 2 rows around 1-8M cells
// (3.41M cells: 2 rows 0.136 ms, 4 rows 0.149), 1 row below.
inline int step_blocks(int64_t n) {
    const int64_t rows = (n + BLOCK - 1) / BLOCK;
    const int64_t run = rows >= 32768 ? 4 : (rows >= 4096 ? 2 : 1);
    return (int)std::max<int64_t>(1, (rows + run - 1) / run);
}

}  // namespace

// ------------------------------------------------------------ engine
struct EngineBase {
    virtual ~EngineBase() = default;
    virtual int init_lattice() = 0;
    virtual int step(int n) = 0;
    virtual int read_field(int id, double* d64, float* d32) = 0;
    virtual int read_populations(double* out) = 0;
    virtual int cv_force(float hx, float hy, float hz, double out[3]) = 0;
    virtual int checksums(double* r, double* u) = 0;
    virtual int thermal_upload(const float* T) = 0;
    virtual int projected_area(int axis, float* out) = 0;
    virtual void scalars(aeros_lbm_scalars* o) = 0;
    virtual int vorticity_q() = 0;
    virtual int sample(int field, const float* pos, int64_t np, float* out, uint8_t* valid) = 0;
    // The same sampler with its device staging (positions, values, flags of np
    // points) in `stage`; allocates nothing, and refuses a field whose device
    // buffer would first have to be allocated.
    virtual int sample_staged(int field, const float* pos, int64_t np, float* out, uint8_t* valid,
                              unsigned char* stage) = 0;
    virtual int sample_velocity(int world, const float* pos, int64_t np, float* out) = 0;
    // A field, as float, into a device buffer on the solver stream (no wait).
    virtual int field_device(int id, float* dst) = 0;
    virtual int step_count() const = 0;   // lbmCurrentStep
    // HIP-event timing of the step kernel launches: 1 = a pair around every
    // launch, 2 = one pair around the launches of each call
    int timing = 0;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
    std::vector<hipEvent_t> events;
};

struct FrameDriver;

struct aeros_lbm {
    aeros_lbm_config user_cfg{};
    std::vector<int32_t> voxels;   // copy of the caller's voxel grid (borrowed only during create)
    Setup s;
    std::unique_ptr<EngineBase> eng;
    hipStream_t stream = nullptr;
    int device = 0;
    mutable std::mutex mu;                  // one call (or frame of the frame worker) at a time
    std::unique_ptr<FrameDriver> frames;    // asynchronous frames, when started
    ~aeros_lbm();
};

namespace {

template <class R, class S, class U> struct Engine final : EngineBase {
    aeros_lbm* h;
    Setup& s;
    hipStream_t st;
    int64_t n = 0;
    int grid = 1;
    DevBuf f[2], mask, solid_u8, u[2], fld[6], world[4], g[2], T[2], buoy[2], parts, acc, status, prev_rho, scratch;
    DevBuf vort, qcrit;   // computeLBMVorticityAndQ outputs, allocated on first use
    bool pre = true;          // buffers hold pre-collision populations (after init / reset)
    int cur = 0;              // lbmCurrentStep (host mirror)
    int resets = 0;
    // post-pass state (aeros/src/lbm.cpp:1662-1705)
    bool prev_valid = false;
    int prev_step = -1;
    float prev_max_u = 0.0f;
    float p_min = 0, p_max = 0, res_rho = 0, res_u = 0;
    int64_t p_bad = 0;
    float nusselt = 0, mean_t = 0.5f;
    float time_ms = 0;
    Status host_status{};

    Engine(aeros_lbm* hh) : h(hh), s(hh->s), st(hh->stream) {}
    ~Engine() override {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }

    int ensure_events(size_t k) {
        while (events.size() < k) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            events.push_back(e);
        }
        return AEROS_OK;
    }

    // ---------------------------------------------------- parameters
    StepParams<R> params() const {
        StepParams<R> P;
        std::memset(&P, 0, sizeof P);
        const aeros_lbm_config& c = s.cfg;
        P.nx = s.nx; P.ny = s.ny; P.nz = s.nz; P.n = n;
        P.axis = s.axis; P.sign = s.sign;
        P.les = c.use_les != 0; P.trt = c.use_trt != 0; P.thermal = (c.thermal && s.thermal_ready) ? 1 : 0;
        P.periodic_x = c.periodic_x != 0; P.periodic_z = c.periodic_z != 0; P.conv_outlet = c.convective_outlet != 0;
        float tau0 = c.tau;   // stepLBMCPU's sanitised tau (aeros/src/lbm.cpp:918-920)
        if (!std::isfinite(tau0) || tau0 < 0.51f) tau0 = 0.55f;
        if (tau0 > 1.5f) tau0 = 1.5f;
        P.tau0 = (R)tau0;
        P.inv_tau0 = R(1) / P.tau0;
        {
            const R mm = P.inv_tau0 - R(0.5);
            const R den = (std::fabs(mm) > R(1e-6)) ? (R(0.1875) / mm + R(0.5)) : R(1);
            P.s_minus0 = R(1) / (den > R(0.05) ? den : R(0.05));
        }
        float cs = c.smagorinsky_c;
        if (!std::isfinite(cs) || cs < 0.0f) cs = 0.12f;
        if (cs > 0.3f) cs = 0.3f;
        P.cs = (R)cs;
        static const double vd[7] = {0.0, 0.3934693403, 0.6321205588, 0.7768698399, 0.8646647168, 0.9179150014, 0.9502129316};
        for (int i = 0; i < 7; ++i) P.vd[i] = (R)vd[i];
        float U0 = std::sqrt(s.inlet[0] * s.inlet[0] + s.inlet[1] * s.inlet[1] + s.inlet[2] * s.inlet[2]);
        if (!std::isfinite(U0) || U0 < 1e-6f) U0 = 0.08f;
        float vs = c.flow_speed / U0;
        if (!std::isfinite(vs) || vs > 1000.0f) vs = 20.0f;
        if (vs < 0.01f) vs = 0.01f;
        P.vel_scale2 = (R)vs * (R)vs;
        P.vel_scale = vs;
        P.fx = (R)c.body_force[0]; P.fy = (R)c.body_force[1]; P.fz = (R)c.body_force[2];
        for (int i = 0; i < 3; ++i) P.in_u[i] = (R)s.inlet[i];
        {   // computeEquilibriumFast(q, 1, u_in) (aeros/src/lbm.cpp:1426-1428)
            const R a = P.in_u[0], b = P.in_u[1], d = P.in_u[2];
            R us = a * a + b * b + d * d;
            R ua = a, ub = b, ud = d;
            if (us > R(0.09)) { const R k = std::sqrt(R(0.09) / us); ua *= k; ub *= k; ud *= k; us = ua * ua + ub * ub + ud * ud; }
            for (int q = 0; q < Q; ++q) {
                R cu = (R)cx_of(q) * ua + (R)cy_of(q) * ub + (R)cz_of(q) * ud;
                if (!std::isfinite(cu)) cu = R(0);
                R fe = w_of<R>(q) * R(1) * (R(1) + cu * R(3) + cu * cu * R(4.5) - us * R(1.5));
                if (!std::isfinite(fe)) fe = w_of<R>(q) * R(1);
                P.feq_in[q] = fe;
            }
            for (int q = 0; q < Q; ++q) P.st_shift[q] = kEqShift<S> ? P.feq_in[q] : w_of<R>(q);
        }
        float cc = U0 * 0.8f;   // aeros/src/lbm.cpp:1467-1469
        if (cc > 0.5f) cc = 0.5f;
        if (cc < 0.01f) cc = 0.08f;
        P.conv_c = (R)cc;
        P.conv_k2 = R(0.15) * (R)cc;
        const int64_t stride[3] = {1, s.nx, (int64_t)s.nx * s.ny};
        P.out_d1 = (int64_t)s.sign * stride[s.axis];
        P.out_d2 = 2 * (int64_t)s.sign * stride[s.axis];
        P.t_inv_tau = R(1) / (R)s.tauT;
        P.t_ref = (R)(0.5f * (c.hot_temp + c.cold_temp));
        P.t_g = (R)s.gravity;
        P.t_hi = (R)(c.hot_temp + 1.0f);
        P.t_lo = (R)(c.cold_temp - 1.0f);
        P.t_hot = (R)c.hot_temp; P.t_cold = (R)c.cold_temp;
        P.t_fixed = c.fixed_wall_temps != 0;
        P.inlet_face = face_bit(s.axis, s.sign > 0 ? 0 : 1);
        P.outlet_face = face_bit(s.axis, s.sign > 0 ? 1 : 0);
        P.wrap_faces = (c.periodic_x ? (M_X0 | M_X1) : 0u) | (c.periodic_z ? (M_Z0 | M_Z1) : 0u);
        P.slip_faces = ((uint32_t)c.free_slip & 63u) << M_FACE_SHIFT;
        P.inv_nx = 1.0 / (double)s.nx; P.half_inv_nx = 0.5 / (double)s.nx;
        P.inv_ny = 1.0 / (double)s.ny; P.half_inv_ny = 0.5 / (double)s.ny;
        P.fluid_cells = s.fluid;
        P.nan_limit = n / 10;
        P.conv_threshold = c.convergence_threshold;
        {   // tau0 is in [0.51, 1.5] and |u| <= 0.3 after the clamps; with a total force
            // below 1e3 every relaxed population stays below ~1e12
            const double F = std::fabs((double)c.body_force[0]) + std::fabs((double)c.body_force[1]) +
                             std::fabs((double)c.body_force[2]) +
                             (P.thermal ? std::fabs((double)P.t_g) * ((double)P.t_hi - (double)P.t_lo) : 0.0);
            P.bounded = std::isfinite(F) && F <= 1e3 && std::isfinite((double)P.t_ref) ? 1 : 0;
        }
        return P;
    }

    // ---------------------------------------------------- setup
    int alloc_all() {
        const size_t NS = (size_t)n;
        int rc;
        for (int b = 0; b < 2; ++b)
            if ((rc = f[b].alloc(NS * Q * sizeof(S)))) return rc;
        if ((rc = mask.alloc(NS * 4))) return rc;
        if ((rc = solid_u8.alloc(NS))) return rc;
        if (s.cfg.use_les)
            for (int b = 0; b < 2; ++b)
                if ((rc = u[b].alloc(NS * VelAos<U>::stride * sizeof(U)))) return rc;
        for (auto& x : fld) if ((rc = x.alloc(NS * sizeof(R)))) return rc;
        for (auto& x : world) if ((rc = x.alloc(NS * sizeof(R)))) return rc;
        if (s.thermal_ready) {
            for (int b = 0; b < 2; ++b) {
                if ((rc = g[b].alloc(NS * Q7 * sizeof(R)))) return rc;
                if ((rc = T[b].alloc(NS * sizeof(R)))) return rc;
                if ((rc = buoy[b].alloc(NS * sizeof(R)))) return rc;
            }
        }
        if ((rc = status.alloc(sizeof(Status)))) return rc;
        grid = step_blocks(n);
        if ((rc = parts.alloc((size_t)2 * grid * sizeof(Partial)))) return rc;
        if ((rc = acc.alloc((size_t)kAccSets * kAccSet * sizeof(int64_t)))) return rc;
        int64_t stride = 1;
        while (n / stride > 262144) stride *= 2;
        if ((rc = prev_rho.alloc((size_t)((n + stride - 1) / stride) * sizeof(float)))) return rc;
        if ((rc = scratch.alloc(std::max<size_t>((size_t)65536 * 3 * sizeof(double), 4096)))) return rc;
        return AEROS_OK;
    }


    int upload_mask() {
        const size_t NS = (size_t)n;
        DevBuf dsol, dgr, dwall;
        int rc;
        if ((rc = dsol.alloc(NS)) || (rc = dgr.alloc(NS)) || (rc = dwall.alloc(NS))) return rc;
        HIP_TRY(hipMemcpyAsync(dsol.p, s.solid.data(), NS, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dgr.p, s.ground.data(), NS, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dwall.p, s.wall.data(), NS, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(solid_u8.p, s.solid.data(), NS, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_build_mask, dim3(grid_for(n)), dim3(256), 0, st, mask.as<uint32_t>(), dsol.as<uint8_t>(),
                           dgr.as<uint8_t>(), dwall.as<uint8_t>(), s.nx, s.ny, s.nz, s.axis, s.sign,
                           (int)(s.cfg.periodic_x != 0), (int)(s.cfg.periodic_z != 0),
                           (int)(s.cfg.convective_outlet != 0), ((uint32_t)s.cfg.free_slip & 63u) << M_FACE_SHIFT);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));   // the temporaries die at return
        return AEROS_OK;
    }

    template <class T> int fill(DevBuf& b, int64_t count, T v) {
        hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(count)), dim3(256), 0, st, b.as<T>(), count, v);
        HIP_TRY(hipGetLastError());
        return AEROS_OK;
    }

    int init_lattice() override {
        n = s.total;
        int rc;
        if ((rc = alloc_all())) return rc;
        if ((rc = upload_mask())) return rc;
        // initial equilibria (aeros/src/lbm.cpp:682-739); curl-noise velocity comes from the host
        DevBuf u3;
        const float* du3 = nullptr;
        if (s.curl_noise) {
            std::vector<float> hu;
            curl_noise_velocity(s, hu);
            if ((rc = u3.alloc(hu.size() * sizeof(float)))) return rc;
            HIP_TRY(hipMemcpyAsync(u3.p, hu.data(), hu.size() * sizeof(float), hipMemcpyHostToDevice, st));
            du3 = u3.as<float>();
        }
        hipLaunchKernelGGL((k_init<R, S>), dim3(grid_for(n)), dim3(256), 0, st, f[0].as<S>(), f[1].as<S>(),
                           mask.as<uint32_t>(), du3, n, (R)s.inlet[0], (R)s.inlet[1], (R)s.inlet[2], params());
        HIP_TRY(hipGetLastError());
        // fields: rho 1, u 0, tke 0, strain 0 (aeros/src/lbm.cpp:567-580)
        if ((rc = fill<R>(fld[0], n, R(1)))) return rc;
        for (int k = 1; k < 6; ++k) if ((rc = fill<R>(fld[k], n, R(0)))) return rc;
        for (auto& x : world) if ((rc = fill<R>(x, n, R(0)))) return rc;
        if (s.cfg.use_les)
            for (int b = 0; b < 2; ++b) if ((rc = fill<U>(u[b], n * VelAos<U>::stride, U(0)))) return rc;
        if (s.thermal_ready) {
            std::vector<float> Th;
            thermal_initial_field(s, Th);
            for (int b = 0; b < 2; ++b) if ((rc = fill<R>(buoy[b], n, R(0)))) return rc;
            if ((rc = upload_T(Th.data()))) return rc;
        }
        if ((rc = fill<int64_t>(acc, (int64_t)kAccSets * kAccSet, 0))) return rc;
        Status z{};
        host_status = z;
        HIP_TRY(hipMemcpyAsync(status.p, &z, sizeof z, hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
        pre = true;
        cur = 0;
        prev_valid = false;
        prev_step = -1;
        thermal_diag();
        return AEROS_OK;
    }

    // T into both parities: the step never writes T of a solid cell, which
    // keeps the uploaded value as the reference's single thermalT array does;
    // populations of the current parity to g_eq(T, 0)
    // (aeros/src/thermal_lbm.cpp:154-197)
    int upload_T(const float* Th) {
        std::vector<R> tr((size_t)n);
        for (int64_t i = 0; i < n; ++i) tr[(size_t)i] = (R)Th[i];
        const int p = cur & 1;
        for (int b = 0; b < 2; ++b)
            HIP_TRY(hipMemcpyAsync(T[b].p, tr.data(), (size_t)n * sizeof(R), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_thermal_eq<R>), dim3(grid_for(n)), dim3(256), 0, st, g[p].as<R>(), T[p].as<R>(), n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
        return AEROS_OK;
    }

    int thermal_upload(const float* Th) override {
        if (!s.thermal_ready) return set_err(AEROS_ERR_STATE, "thermal module not initialised");
        int rc = upload_T(Th);
        if (rc) return rc;
        return thermal_diag();
    }

    int thermal_diag() {   // thermalMeanT and thermalNusselt of the current T
        if (!s.thermal_ready) return AEROS_OK;
        const int p = cur & 1;
        const int64_t m = (int64_t)(s.yhi - s.ylo + 1) * s.nx * s.nz;
        const int gb = (int)std::max<int64_t>(1, std::min<int64_t>((m + BLOCK - 1) / BLOCK, 4096));
        hipLaunchKernelGGL((k_thermal_diag<R>), dim3(gb), dim3(BLOCK), 0, st, T[p].as<R>(), mask.as<uint32_t>(), s.nx,
                           s.ny, s.nz, s.ylo, s.yhi, (R)s.cfg.hot_temp, scratch.as<double>());
        HIP_TRY(hipGetLastError());
        std::vector<double> hp((size_t)gb * 3);
        HIP_TRY(hipMemcpyAsync(hp.data(), scratch.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        double a = 0, b = 0, c = 0;
        for (int k = 0; k < gb; ++k) { a += hp[k * 3]; b += hp[k * 3 + 1]; c += hp[k * 3 + 2]; }
        const long cnt = (long)s.nx * (long)s.nz * (long)(s.yhi - s.ylo + 1);
        mean_t = cnt > 0 ? (float)(a / (double)cnt) : 0.5f;
        if (cur > 0 && c > 0) {   // Nu is produced by a thermal step
            const float dT = std::max(1e-6f, std::fabs(s.cfg.hot_temp - s.cfg.cold_temp));
            nusselt = (float)(b / c) * s.H / dT;
        }
        return AEROS_OK;
    }

    // ---------------------------------------------------- stepping
    StepBuffers buffers(int k, bool first, bool fields) {
        StepBuffers B{};
        const int p = k & 1;   // parity of reference step k+1's output
        B.f_in = first ? f[p].p : f[p ^ 1].p;
        B.f_out = f[p].p;
        B.mask = mask.as<uint32_t>();
        if (s.cfg.use_les) { B.u_prev = u[p].p; B.u_next = u[p ^ 1].p; }
        for (int c = 0; c < 6; ++c) B.fld[c] = fld[c].p;
        if (s.thermal_ready) {
            B.g_in = g[p].p; B.g_out = g[p ^ 1].p;
            B.b_in = buoy[p].p; B.b_out = buoy[p ^ 1].p;
            B.t_out = T[p ^ 1].p;
        }
        B.parts = parts.as<Partial>();
        B.acc = acc.as<int64_t>();
        B.status = status.as<Status>();
        (void)fields;
        return B;
    }

    int step(int nsteps) override {
        if (nsteps <= 0) return AEROS_OK;
        const auto t0 = std::chrono::steady_clock::now();
        StepParams<R> P = params();
        const int start = cur;
        const bool was_pre = pre;
        if (timing) {
            const int rc = ensure_events(timing == 1 ? (size_t)nsteps * 2 : 2);
            if (rc) return rc;
        }
        if (timing == 2) HIP_TRY(hipEventRecord(events[0], st));
        for (int i = 0; i < nsteps; ++i) {
            const int k = start + i;
            P.first = (was_pre && i == 0) ? 1 : 0;
            P.write_fields = (i == nsteps - 1) ? 1 : 0;
            P.dry = 0;
            P.k = k;
            P.pending = i > 0 ? 1 : 0;
            const StepBuffers B = buffers(k, P.first != 0, P.write_fields != 0);
            if (timing == 1) HIP_TRY(hipEventRecord(events[(size_t)i * 2], st));
            HIP_TRY((launch_step<R, S, U>(P, B, grid, st)));
            if (timing == 1) HIP_TRY(hipEventRecord(events[(size_t)i * 2 + 1], st));
        }
        if (timing == 2) HIP_TRY(hipEventRecord(events[1], st));
        {   // fold the call's last step into the status
            const int k = start + nsteps - 1;
            hipLaunchKernelGGL((k_fold<R>), dim3(1), dim3(BLOCK), 0, st, P, buffers(k, false, false), k, (uint32_t)grid);
            HIP_TRY(hipGetLastError());
        }
        int rc = finish(start, nsteps, was_pre);
        if (rc) return rc;
        if (timing) {
            for (int i = 0; i < (timing == 1 ? nsteps : 1); ++i) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, events[(size_t)i * 2], events[(size_t)i * 2 + 1]));
                timed_ms += ms;
            }
            timed_launches += nsteps;
        }
        rc = post_pass();
        if (rc) return rc;
        rc = thermal_diag();
        if (rc) return rc;
        time_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return AEROS_OK;
    }

    // After nsteps launches from step index `start`: read the status and, if
    // the loop broke, finish the broken step and reset the fluid as the
    // reference does.  Shared by single-lattice stepping and served rounds.
    int finish(int start, int nsteps, bool was_pre) {
        Status S_{};
        HIP_TRY(hipMemcpyAsync(&S_, status.p, sizeof S_, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return finish_status(S_, start, nsteps, was_pre);
    }
    // The same with the status already on the host (served mode reads every
    // case's status with one copy).
    int finish_status(Status S_, int start, int nsteps, bool was_pre) {
        if (S_.abort == 0) {
            cur = start + nsteps;
            pre = false;
        } else {
            StepParams<R> P = params();
            // the loop broke at reference step s (aeros/src/lbm.cpp:1313-1319, 1574-1580)
            const int sidx = S_.abort_step;
            const int i = sidx - start - 1;              // index of that kernel within this call
            if (i != nsteps - 1) {                       // its fields were not written: re-run it
                P.first = (was_pre && i == 0) ? 1 : 0;
                P.write_fields = 1;
                P.dry = 1;
                P.k = sidx - 1;
                P.pending = 0;
                const StepBuffers B = buffers(sidx - 1, P.first != 0, true);
                HIP_TRY((launch_step<R, S, U>(P, B, grid, st)));
            }
            cur = (S_.abort == 1) ? sidx - 1 : sidx;
            // reset the fluid to the inflow equilibrium (aeros/src/lbm.cpp:1583-1600)
            const int p = cur & 1;
            hipLaunchKernelGGL((k_reset<R, S, U>), dim3(grid_for(n)), dim3(256), 0, st, f[p].as<S>(),
                               mask.as<uint32_t>(), n, P, s.cfg.use_les ? u[p].as<U>() : nullptr,
                               fld[0].as<R>(), fld[1].as<R>(), fld[2].as<R>(), fld[3].as<R>());
            HIP_TRY(hipGetLastError());
            int rc2 = fill<int64_t>(acc, (int64_t)kAccSets * kAccSet, 0);   // step indices restart from cur
            if (rc2) return rc2;
            pre = true;
            resets++;
            S_.abort = 0;
            S_.abort_step = 0;
            S_.avg_rho = 1.0f;
            const float a = s.inlet[0], b = s.inlet[1], c = s.inlet[2];
            S_.avg_ke = (float)((R)a * (R)a + (R)b * (R)b + (R)c * (R)c);
            HIP_TRY(hipMemcpyAsync(status.p, &S_, sizeof S_, hipMemcpyHostToDevice, st));
        }
        host_status = S_;
        return AEROS_OK;
    }

    // Served mode: this lattice's entry of a batched launch, on nblocks
    // blocks (at most `grid`, the size of the partials buffer).
    void batch_entry(BatchCase<R>& bc, uint32_t nblocks) {
        bc.P = params();
        bc.P.first = bc.P.write_fields = bc.P.dry = 0;
        bc.buf[0] = buffers(0, false, false);
        bc.buf[1] = buffers(1, false, false);
        bc.start = cur;
        bc.pre = pre ? 1 : 0;
        bc.nblocks = std::min<uint32_t>(nblocks, (uint32_t)grid);
        bc.pad = 0;
    }
    int64_t device_bytes() const {
        int64_t b = 0;
        for (const DevBuf* d : {&f[0], &f[1], &mask, &solid_u8, &u[0], &u[1], &g[0], &g[1], &T[0], &T[1], &buoy[0],
                                &buoy[1], &parts, &status, &prev_rho, &scratch})
            b += (int64_t)d->bytes;
        for (const DevBuf& d : fld) b += (int64_t)d.bytes;
        for (const DevBuf& d : world) b += (int64_t)d.bytes;
        return b;
    }

    // world velocity, pressure and the sampled residuals (aeros/src/lbm.cpp:1602-1705)
    int post_pass() {
        float U0 = std::sqrt(s.inlet[0] * s.inlet[0] + s.inlet[1] * s.inlet[1] + s.inlet[2] * s.inlet[2]);
        if (!std::isfinite(U0) || U0 < 1e-6f) U0 = 0.08f;
        float scale = s.cfg.flow_speed / U0;
        if (!std::isfinite(scale) || scale > 1000.0f) scale = 20.0f;
        unsigned long long* bad = reinterpret_cast<unsigned long long*>(scratch.p);
        HIP_TRY(hipMemsetAsync(bad, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL((k_post<R>), dim3(grid_for(n)), dim3(256), 0, st, mask.as<uint32_t>(), fld[0].as<R>(),
                           fld[1].as<R>(), fld[2].as<R>(), fld[3].as<R>(), world[0].as<R>(), world[1].as<R>(),
                           world[2].as<R>(), world[3].as<R>(), n, (R)scale, (R)s.cfg.air_density, bad);
        HIP_TRY(hipGetLastError());
        int64_t stride = 1;
        while (n / stride > 262144) stride *= 2;
        const bool valid = prev_valid && prev_step >= 0 && cur >= prev_step;
        float* out = reinterpret_cast<float*>(scratch.as<char>() + 64);
        PostPart* pparts = reinterpret_cast<PostPart*>(scratch.as<char>() + 4096);
        const int64_t ns = (n + stride - 1) / stride;
        const int pgrid = (int)std::max<int64_t>(1, std::min<int64_t>((ns + BLOCK - 1) / BLOCK, kPostBlocks));
        hipLaunchKernelGGL((k_post_scalars<R>), dim3(pgrid), dim3(BLOCK), 0, st, world[3].as<R>(), fld[0].as<R>(),
                           prev_rho.as<float>(), n, stride, (int)valid, pparts);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_post_fold, dim3(1), dim3(BLOCK), 0, st, pparts, pgrid, (int)valid, out);
        HIP_TRY(hipGetLastError());
        struct { unsigned long long bad; char pad[56]; float v[4]; } hb;
        HIP_TRY(hipMemcpyAsync(&hb, scratch.p, sizeof hb, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        p_bad = (int64_t)hb.bad;
        p_min = hb.v[0]; p_max = hb.v[1]; res_rho = hb.v[2];
        res_u = std::fabs(host_status.max_u_lb - prev_max_u);
        prev_max_u = host_status.max_u_lb;
        prev_valid = true;
        prev_step = cur;
        return AEROS_OK;
    }

    // ---------------------------------------------------- readback
    int read_field(int id, double* d64, float* d32) override {
        const DevBuf* src = nullptr;
        switch (id) {
            case AEROS_FIELD_RHO: src = &fld[0]; break;
            case AEROS_FIELD_UX: src = &fld[1]; break;
            case AEROS_FIELD_UY: src = &fld[2]; break;
            case AEROS_FIELD_UZ: src = &fld[3]; break;
            case AEROS_FIELD_TKE: src = &fld[4]; break;
            case AEROS_FIELD_STRAIN: src = &fld[5]; break;
            case AEROS_FIELD_UX_WORLD: src = &world[0]; break;
            case AEROS_FIELD_UY_WORLD: src = &world[1]; break;
            case AEROS_FIELD_UZ_WORLD: src = &world[2]; break;
            case AEROS_FIELD_PRESSURE: src = &world[3]; break;
            case AEROS_FIELD_VORTICITY:
            case AEROS_FIELD_QCRIT: {
                const DevBuf& b = id == AEROS_FIELD_VORTICITY ? vort : qcrit;
                if (!b.p) {   // never computed: zeros, as the reference's freshly allocated arrays
                    for (int64_t i = 0; i < n; ++i) { if (d64) d64[i] = 0.0; if (d32) d32[i] = 0.0f; }
                    return AEROS_OK;
                }
                src = &b;
                break;
            }
            case AEROS_FIELD_T:
            case AEROS_FIELD_BUOYANCY:
                if (!s.thermal_ready) return set_err(AEROS_ERR_STATE, "thermal module is off");
                src = id == AEROS_FIELD_T ? &T[cur & 1] : &buoy[cur & 1];
                break;
            case AEROS_FIELD_SOLID:
            case AEROS_FIELD_GROUND:
            case AEROS_FIELD_WALLDIST: {
                const std::vector<uint8_t>& v = id == AEROS_FIELD_SOLID ? s.solid : (id == AEROS_FIELD_GROUND ? s.ground : s.wall);
                for (int64_t i = 0; i < n; ++i) {
                    if (d64) d64[i] = v[(size_t)i];
                    if (d32) d32[i] = v[(size_t)i];
                }
                return AEROS_OK;
            }
            default: return set_err(AEROS_ERR_ARG, "unknown field %d", id);
        }
        std::vector<R> tmp((size_t)n);
        HIP_TRY(hipMemcpyAsync(tmp.data(), src->p, (size_t)n * sizeof(R), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int64_t i = 0; i < n; ++i) {
            if (d64) d64[i] = (double)tmp[(size_t)i];
            if (d32) d32[i] = (float)tmp[(size_t)i];
        }
        return AEROS_OK;
    }

    int field_device(int id, float* dst) override {
        const DevBuf* src = nullptr;
        switch (id) {
            case AEROS_FIELD_RHO: src = &fld[0]; break;
            case AEROS_FIELD_UX: src = &fld[1]; break;
            case AEROS_FIELD_UY: src = &fld[2]; break;
            case AEROS_FIELD_UZ: src = &fld[3]; break;
            case AEROS_FIELD_TKE: src = &fld[4]; break;
            case AEROS_FIELD_STRAIN: src = &fld[5]; break;
            case AEROS_FIELD_UX_WORLD: src = &world[0]; break;
            case AEROS_FIELD_UY_WORLD: src = &world[1]; break;
            case AEROS_FIELD_UZ_WORLD: src = &world[2]; break;
            case AEROS_FIELD_PRESSURE: src = &world[3]; break;
            case AEROS_FIELD_VORTICITY: src = vort.p ? &vort : nullptr; break;
            case AEROS_FIELD_QCRIT: src = qcrit.p ? &qcrit : nullptr; break;
            case AEROS_FIELD_T:
            case AEROS_FIELD_BUOYANCY:
                if (!s.thermal_ready) return set_err(AEROS_ERR_STATE, "thermal module is off");
                src = id == AEROS_FIELD_T ? &T[cur & 1] : &buoy[cur & 1];
                break;
            default: return set_err(AEROS_ERR_ARG, "field %d is not held on the device", id);
        }
        if (!src) {   // vorticity / Q never computed: zeros, as read_field gives
            hipLaunchKernelGGL((k_fill<float>), dim3(grid_for(n)), dim3(256), 0, st, dst, n, 0.0f);
        } else if constexpr (sizeof(R) == sizeof(float)) {
            HIP_TRY(hipMemcpyAsync(dst, src->p, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL((k_to_f32<R>), dim3(grid_for(n)), dim3(256), 0, st, src->as<R>(), dst, n);
        }
        HIP_TRY(hipGetLastError());
        return AEROS_OK;
    }

    int step_count() const override { return cur; }

    int read_populations(double* out) override {
        DevBuf tmp;
        int rc = tmp.alloc((size_t)n * Q * sizeof(R));
        if (rc) return rc;
        const int p = cur & 1;
        if (pre) {
            hipLaunchKernelGGL((k_widen<R, S>), dim3(grid_for(n)), dim3(256), 0, st, f[p].as<S>(), tmp.as<R>(), n,
                               params());
        } else {
            StepParams<R> P = params();
            hipLaunchKernelGGL((k_pull_only<R, S>), dim3(grid_for(n)), dim3(256), 0, st, f[p ^ 1].as<S>(), f[p].as<S>(),
                               tmp.as<R>(), mask.as<uint32_t>(), P);
        }
        HIP_TRY(hipGetLastError());
        std::vector<R> h_((size_t)n * Q);
        HIP_TRY(hipMemcpyAsync(h_.data(), tmp.p, h_.size() * sizeof(R), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int64_t c = 0; c < n; ++c)
            for (int q = 0; q < Q; ++q) out[c * Q + q] = (double)h_[(size_t)q * n + c];
        return AEROS_OK;
    }

    // computeLBMVorticityAndQ (aeros/src/lbm.cpp:2005-2093) on the device
    int vorticity_q() override {
        int rc;
        if (!vort.p) {
            if ((rc = vort.alloc((size_t)n * sizeof(R))) || (rc = qcrit.alloc((size_t)n * sizeof(R)))) return rc;
            if ((rc = fill<R>(vort, n, R(0))) || (rc = fill<R>(qcrit, n, R(0)))) return rc;
        }
        float cs[3];
        for (int k = 0; k < 3; ++k) cs[k] = s.cell[k] < 1e-6f ? 0.1f : s.cell[k];
        hipLaunchKernelGGL((k_vort_q<R>), dim3(grid_for(n)), dim3(256), 0, st, mask.as<uint32_t>(), world[0].as<R>(),
                           world[1].as<R>(), world[2].as<R>(), vort.as<R>(), qcrit.as<R>(), fld[5].as<R>(), s.nx, s.ny,
                           s.nz, 1.0f / cs[0], 1.0f / cs[1], 1.0f / cs[2]);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
        return AEROS_OK;
    }

    SampleGrid sample_grid() const {
        SampleGrid G;
        G.minx = s.min[0]; G.miny = s.min[1]; G.minz = s.min[2];
        G.csx = s.cell[0]; G.csy = s.cell[1]; G.csz = s.cell[2];
        G.maxx = s.min[0] + (float)s.nx * s.cell[0];
        G.maxy = s.min[1] + (float)s.ny * s.cell[1];
        G.maxz = s.min[2] + (float)s.nz * s.cell[2];
        G.nx = s.nx; G.ny = s.ny; G.nz = s.nz;
        return G;
    }

    // The field and mode of a scalar sampler; with may_alloc false a field whose
    // device buffer does not exist yet (vorticity, Q) is refused.
    int sample_source(int field, bool may_alloc, const R** src, int* mode) {
        *src = nullptr;
        *mode = 0;
        switch (field) {
            case AEROS_FIELD_PRESSURE: *src = world[3].as<R>(); break;
            case AEROS_FIELD_RHO: *src = fld[0].as<R>(); *mode = 1; break;
            case AEROS_FIELD_TKE: *src = fld[4].as<R>(); break;
            case AEROS_FIELD_STRAIN: *src = fld[5].as<R>(); break;
            case AEROS_FIELD_VORTICITY:
            case AEROS_FIELD_QCRIT: {
                if (!vort.p) {
                    if (!may_alloc)
                        return set_err(AEROS_ERR_STATE, "vorticity and Q are not on the device yet: call "
                                                        "aeros_lbm_vorticity_q first");
                    if (int rc = vorticity_q()) return rc;
                }
                *src = (field == AEROS_FIELD_VORTICITY ? vort : qcrit).as<R>();
                break;
            }
            default: return set_err(AEROS_ERR_ARG, "field %d has no world sampler", field);
        }
        return AEROS_OK;
    }

    // Positions in, values and flags out through device staging dp / dout / dval.
    int sample_run(const R* src, int mode, const float* pos, int64_t np, float* out, uint8_t* valid, float* dp,
                   float* dout, uint8_t* dval) {
        HIP_TRY(hipMemcpyAsync(dp, pos, (size_t)np * 3 * sizeof(float), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_sample<R>), dim3(grid_for(np)), dim3(256), 0, st, src, mask.as<uint32_t>(), sample_grid(),
                           dp, np, mode, dout, dval);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, dout, (size_t)np * sizeof(float), hipMemcpyDeviceToHost, st));
        if (valid) HIP_TRY(hipMemcpyAsync(valid, dval, (size_t)np, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return AEROS_OK;
    }

    // Batched world-space samplers (aeros/src/lbm.cpp:1845-1999) on the device.
    int sample(int field, const float* pos, int64_t np, float* out, uint8_t* valid) override {
        const R* src;
        int mode;
        if (int rc = sample_source(field, true, &src, &mode)) return rc;
        if (np <= 0) return AEROS_OK;
        DevBuf dp, dout, dval;
        int rc;
        if ((rc = dp.alloc((size_t)np * 3 * sizeof(float))) || (rc = dout.alloc((size_t)np * sizeof(float))) ||
            (rc = dval.alloc((size_t)np)))
            return rc;
        return sample_run(src, mode, pos, np, out, valid, dp.as<float>(), dout.as<float>(), dval.as<uint8_t>());
    }

    // stage: 3 np floats of positions, then np floats of values, then np flags
    int sample_staged(int field, const float* pos, int64_t np, float* out, uint8_t* valid,
                      unsigned char* stage) override {
        const R* src;
        int mode;
        if (int rc = sample_source(field, false, &src, &mode)) return rc;
        if (np <= 0) return AEROS_OK;
        float* dp = reinterpret_cast<float*>(stage);
        float* dout = dp + 3 * np;
        return sample_run(src, mode, pos, np, out, valid, dp, dout, reinterpret_cast<uint8_t*>(dout + np));
    }

    int sample_velocity(int world_units, const float* pos, int64_t np, float* out) override {
        if (np <= 0) return AEROS_OK;
        float in[3];
        inlet_velocity(s.cfg, in[0], in[1], in[2]);
        float U0 = std::sqrt(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
        if (!std::isfinite(U0) || U0 < 1e-6f) U0 = 0.08f;
        float scale = s.cfg.flow_speed / U0;
        if (!std::isfinite(scale) || scale > 1000.0f) scale = 20.0f;
        float fs[3];
        free_stream(s.cfg, fs);
        DevBuf dp, dout;
        int rc;
        if ((rc = dp.alloc((size_t)np * 3 * sizeof(float))) || (rc = dout.alloc((size_t)np * 3 * sizeof(float)))) return rc;
        HIP_TRY(hipMemcpyAsync(dp.p, pos, (size_t)np * 3 * sizeof(float), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_sample_velocity<R>), dim3(grid_for(np)), dim3(256), 0, st, fld[1].as<R>(), fld[2].as<R>(),
                           fld[3].as<R>(), mask.as<uint32_t>(), sample_grid(), dp.as<float>(), np, world_units ? 1 : 0,
                           scale, s.cfg.flow_speed * 2.0f, fs[0], fs[1], fs[2], dout.as<float>());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)np * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return AEROS_OK;
    }

    // lbmControlVolumeForce box (aeros/src/lbm.cpp:2188-2198)
    // The control volume of lbmControlVolumeForce (aeros/src/lbm.cpp:2183-2200):
    // half extents as fractions of the domain around its centre, kept one
    // cell inside the shell; with the fields it reads.
    CvCase<R> cv_case(float hxf, float hyf, float hzf) const {
        auto side = [](int v, int N) { return v < 1 ? 1 : (v > N - 2 ? N - 2 : v); };
        const int nx = s.nx, ny = s.ny, nz = s.nz;
        const int cx = (int)(nx * 0.5f), cy = (int)(ny * 0.5f), cz = (int)(nz * 0.5f);
        const int hx = std::max(2, (int)(nx * hxf)), hy = std::max(2, (int)(ny * hyf)), hz = std::max(2, (int)(nz * hzf));
        CvCase<R> c;
        c.solid = solid_u8.as<uint8_t>();
        c.rho = fld[0].as<R>(); c.ux = fld[1].as<R>(); c.uy = fld[2].as<R>(); c.uz = fld[3].as<R>();
        c.nx = nx; c.ny = ny; c.nz = nz;
        c.x0 = side(cx - hx, nx); c.x1 = side(cx + hx, nx); c.y0 = side(cy - hy, ny); c.y1 = side(cy + hy, ny);
        c.z0 = side(cz - hz, nz); c.z1 = side(cz + hz, nz);
        return c;
    }

    int cv_force(float hxf, float hyf, float hzf, double out[3]) override {
        const CvCase<R> c = cv_case(hxf, hyf, hzf);
        const int gb = kCvBlocks;
        hipLaunchKernelGGL((k_cv_force<R>), dim3(gb), dim3(BLOCK), 0, st, c.solid, c.rho, c.ux, c.uy, c.uz, c.nx, c.ny,
                           c.nz, c.x0, c.x1, c.y0, c.y1, c.z0, c.z1, scratch.as<double>());
        HIP_TRY(hipGetLastError());
        std::vector<double> hp((size_t)gb * 3);
        HIP_TRY(hipMemcpyAsync(hp.data(), scratch.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        double fx = 0, fy = 0, fz = 0;
        for (int k = 0; k < gb; ++k) { fx += hp[k * 3]; fy += hp[k * 3 + 1]; fz += hp[k * 3 + 2]; }
        out[0] = -fx; out[1] = -fy; out[2] = -fz;
        return AEROS_OK;
    }

    int checksums(double* r, double* uu) override {
        const int gb = (int)std::min<int64_t>((n + BLOCK - 1) / BLOCK, 1024);
        const bool exact = sizeof(R) == 4;
        unsigned long long* ex = reinterpret_cast<unsigned long long*>(scratch.as<char>() + (size_t)gb * 2 * sizeof(double));
        HIP_TRY(hipMemsetAsync(ex, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL((k_checksums<R>), dim3(gb), dim3(BLOCK), 0, st, fld[0].as<R>(), fld[1].as<R>(), fld[2].as<R>(),
                           fld[3].as<R>(), n, scratch.as<double>(), ex, (int)exact);
        HIP_TRY(hipGetLastError());
        std::vector<double> hp((size_t)gb * 2);
        unsigned long long e = 0;
        HIP_TRY(hipMemcpyAsync(hp.data(), scratch.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&e, ex, sizeof e, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        double a = 0, b = 0;
        for (int k = 0; k < gb; ++k) { a += hp[k * 2]; b += hp[k * 2 + 1]; }
        *r = exact ? (double)e / 16777216.0 : a;
        *uu = b;
        return AEROS_OK;
    }

    int projected_area(int axis, float* out) override {
        unsigned long long* cnt = reinterpret_cast<unsigned long long*>(scratch.p);
        HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_proj_area, dim3(256), dim3(256), 0, st, solid_u8.as<uint8_t>(), s.nx, s.ny, s.nz, axis, cnt);
        HIP_TRY(hipGetLastError());
        unsigned long long c = 0;
        HIP_TRY(hipMemcpyAsync(&c, cnt, sizeof c, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *out = (float)(double)c;
        return AEROS_OK;
    }

    void scalars(aeros_lbm_scalars* o) override {
        std::memset(o, 0, sizeof *o);
        const Status& S_ = host_status;
        o->step = cur; o->converged = S_.converged; o->resets = resets;
        o->nx = s.nx; o->ny = s.ny; o->nz = s.nz;
        o->avg_rho = cur == 0 && resets == 0 ? 1.0f : S_.avg_rho;
        o->avg_ke = S_.avg_ke; o->max_u_lb = S_.max_u_lb; o->max_u_world = S_.max_u_world; o->tke = S_.tke;
        o->convergence = S_.convergence;
        o->neg_pops = S_.neg; o->bad_pops = S_.bad; o->rho_out_of_range = S_.rho_out; o->rho_non_finite = S_.rho_nf;
        o->u_limited = S_.ulim; o->pressure_non_finite = p_bad;
        o->tau = s.tau; o->viscosity = s.viscosity; o->U0 = s.U0; o->reynolds_phys = s.re_phys; o->reynolds_lb = s.re_lb;
        o->blockage = s.blockage; o->pressure_min = p_min; o->pressure_max = p_max;
        o->residual_rho = res_rho; o->residual_u = res_u; o->time_ms = time_ms; o->fluid_cells = s.fluid;
        o->nusselt = nusselt; o->thermal_mean_t = mean_t; o->thermal_alpha = s.alpha; o->thermal_gravity = s.gravity;
        o->thermal_tau = s.tauT; o->thermal_rayleigh = s.rayleigh; o->thermal_prandtl = s.prandtl;
        o->thermal_ready = s.thermal_ready; o->thermal_ylo = s.ylo; o->thermal_yhi = s.yhi;
    }
};

EngineBase* make_engine(aeros_lbm* h) {
    switch (h->s.cfg.precision) {
        case AEROS_FP32: return new (std::nothrow) Engine<float, float, float>(h);
        case AEROS_FP64: return new (std::nothrow) Engine<double, double, double>(h);
        case AEROS_FP16: return new (std::nothrow) Engine<float, _Float16, _Float16>(h);
        case AEROS_BF16: return new (std::nothrow) Engine<float, bf16_t, _Float16>(h);
        case AEROS_FP32_FP64C: return new (std::nothrow) Engine<double, float, float>(h);
        case AEROS_FP16_EQ: return new (std::nothrow) Engine<float, h16e_t, _Float16>(h);
        case AEROS_BF16_EQ: return new (std::nothrow) Engine<float, b16e_t, _Float16>(h);
        default: return nullptr;
    }
}

int check_device(int dev) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return set_err(AEROS_ERR_NO_DEVICE, "no HIP device available");
    }
    if (dev < 0 || dev >= count) return set_err(AEROS_ERR_NO_DEVICE, "device %d out of range (%d devices)", dev, count);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(AEROS_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", dev, prop.gcnArchName);
    HIP_TRY(hipSetDevice(dev));
    return AEROS_OK;
}

int build_handle(aeros_lbm* h) {
    char msg[512] = {0};
    int rc = setup_domain(h->user_cfg, h->s, msg, sizeof msg);
    if (rc) return set_err(rc, "%s", msg);
    // kernels index with 32-bit offsets: 19 n elements (double) or 19 n sizeof(S)
    // bytes (float, half, bf16) must stay below 2^31 / 2^32
    {
        const int prec = h->s.cfg.precision;
        const int64_t es = prec == AEROS_FP64 ? 8 : ((prec == AEROS_FP32 || prec == AEROS_FP32_FP64C) ? 4 : 2);
        const int64_t lim = prec == AEROS_FP64 ? (((int64_t)1 << 31) / Q - 1) : (((int64_t)1 << 32) / (Q * es) - 1);
        if (h->s.total > lim)
            return set_err(AEROS_ERR_TOO_LARGE, "%lld cells exceed the %lld-cell limit of one lattice at this precision",
                           (long long)h->s.total, (long long)lim);
    }
    rc = setup_after_mask(h->s, msg, sizeof msg);
    if (rc) return set_err(rc, "%s", msg);
    h->eng.reset(make_engine(h));
    if (!h->eng) return set_err(AEROS_ERR_UNSUPPORTED, "precision %d not built", h->s.cfg.precision);
    rc = h->eng->init_lattice();
    if (rc) h->eng.reset();
    return rc;
}

}  // namespace

// Device staging of aeros_lbm_sample_noalloc, the sampler for callers that
// cannot allocate: positions (3 floats), value (float) and flag (byte) of up
// to kNoAllocPoints points.  Each device that runs the library holds its own
// copy until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host it is only a stand-in for the symbol): it uses
// the device address hipGetSymbolAddress returns, taken once per device on the
// first call there with that device current.
constexpr int64_t kNoAllocPoints = 65536;
constexpr size_t kNoAllocBytes = (size_t)kNoAllocPoints * (3 * sizeof(float) + sizeof(float) + 1);
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocBytes];

namespace {
std::mutex g_noalloc_mu;                                   // guards the address table
unsigned char* g_noalloc_ptr[kNoAllocDevices] = {};        // device address of noalloc_scratch per device
std::mutex g_noalloc_use[kNoAllocDevices];                 // one call at a time per device

// The current device's address of noalloc_scratch (the caller has made `dev` current).
int noalloc_stage(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kNoAllocDevices)
        return set_err(AEROS_ERR_UNSUPPORTED, "device %d is beyond the %d devices of the allocation-free staging",
                       dev, kNoAllocDevices);
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return set_err(AEROS_ERR_HIP, "hipGetSymbolAddress of the sampler staging failed on device %d: %s", dev,
                           hipGetErrorString(e));
        }
        g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return AEROS_OK;
}
}  // namespace

// Serialises a call with the handle's other calls and its frame worker; a
// handle whose last reset or set-mask failed holds no lattice and refuses
// every call but those two.
#define AEROS_LOCK_HANDLE(h)                               \
    std::lock_guard<std::mutex> lock_((h)->mu);            \
    if (!(h)->eng) return set_err(AEROS_ERR_STATE, "the handle holds no lattice (its last reset or set-mask failed)")

namespace {

}  // namespace

#include "lbm_frames.inc"

// ============================================================================ C ABI
extern "C" {

void aeros_lbm_config_defaults(aeros_lbm_config* cfg) {
    if (cfg) config_defaults(cfg);
}

aeros_lbm* aeros_lbm_create(const aeros_lbm_config* cfg, int* err) {
    auto fail = [&](int code) -> aeros_lbm* { if (err) *err = code; return nullptr; };
    if (!cfg) return fail(set_err(AEROS_ERR_ARG, "null config"));
    if (cfg->struct_size != (int32_t)sizeof(aeros_lbm_config))
        return fail(set_err(AEROS_ERR_ARG, "config struct_size %d != %d", cfg->struct_size, (int)sizeof(aeros_lbm_config)));
    if (cfg->precision < AEROS_FP32 || cfg->precision > AEROS_BF16_EQ)
        return fail(set_err(AEROS_ERR_ARG, "unknown precision %d", cfg->precision));
    {
        char msg[160];
        if (check_boundaries(*cfg, msg, (int)sizeof msg)) return fail(set_err(AEROS_ERR_ARG, "%s", msg));
    }
    int rc = check_device(cfg->device);
    if (rc) return fail(rc);
    std::unique_ptr<aeros_lbm> h(new (std::nothrow) aeros_lbm());
    if (!h) return fail(set_err(AEROS_ERR_NOMEM, "handle allocation failed"));
    h->user_cfg = *cfg;
    if (cfg->voxels && cfg->vox_n[0] > 0 && cfg->vox_n[1] > 0 && cfg->vox_n[2] > 0) {
        const size_t nv = (size_t)cfg->vox_n[0] * cfg->vox_n[1] * cfg->vox_n[2];
        h->voxels.assign(cfg->voxels, cfg->voxels + nv);
        h->user_cfg.voxels = h->voxels.data();
    } else {
        h->user_cfg.voxels = nullptr;
    }
    h->device = cfg->device;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_err(AEROS_ERR_HIP, "hipStreamCreate failed"));
    }
    rc = build_handle(h.get());
    if (rc) {
        h->eng.reset();
        (void)hipStreamDestroy(h->stream);
        return fail(rc);
    }
    if (err) *err = AEROS_OK;
    return h.release();
}

int aeros_lbm_destroy(aeros_lbm* h) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    h->frames.reset();   // joins the frame worker
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->eng.reset();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return AEROS_OK;
}

int aeros_lbm_reset(aeros_lbm* h) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    std::lock_guard<std::mutex> lock_(h->mu);
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    h->eng.reset();
    return build_handle(h);   // on failure the handle keeps no lattice and refuses other calls
}

int aeros_lbm_dims(const aeros_lbm* h, int32_t dims[3], float box_min[3], float cell_size[3]) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);
    if (dims) { dims[0] = h->s.nx; dims[1] = h->s.ny; dims[2] = h->s.nz; }
    for (int i = 0; i < 3; ++i) {
        if (box_min) box_min[i] = h->s.min[i];
        if (cell_size) cell_size[i] = h->s.cell[i];
    }
    return AEROS_OK;
}

int aeros_lbm_set_mask(aeros_lbm* h, const uint8_t* solid, int32_t nx, int32_t ny, int32_t nz) {
    if (!h || !solid) return set_err(AEROS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lock_(h->mu);
    if (nx != h->s.nx || ny != h->s.ny || nz != h->s.nz)
        return set_err(AEROS_ERR_ARG, "mask %dx%dx%d does not match the lattice %dx%dx%d", nx, ny, nz, h->s.nx, h->s.ny, h->s.nz);
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    // The new setup is built aside and replaces the handle's only when it is
    // complete: the mask, then the ground slab on top of it as initLBM
    // applies it after the mask transfer (aeros/src/lbm.cpp:596-643), then
    // wall distance, fluid count and tau.
    Setup ns = h->s;
    ns.cfg = h->user_cfg;
    for (int64_t i = 0; i < ns.total; ++i) { ns.solid[(size_t)i] = solid[i] ? 1 : 0; ns.ground[(size_t)i] = 0; }
    apply_ground(ns);
    char msg[256] = {0};
    int rc = setup_after_mask(ns, msg, sizeof msg);
    if (rc) return set_err(rc, "%s", msg);
    h->eng.reset();
    h->s = std::move(ns);
    h->eng.reset(make_engine(h));
    if (!h->eng) return set_err(AEROS_ERR_UNSUPPORTED, "precision not built");
    rc = h->eng->init_lattice();
    if (rc) h->eng.reset();   // no half-built lattice: the handle refuses calls until a reset or set-mask succeeds
    return rc;
}

int aeros_lbm_step(aeros_lbm* h, int32_t steps) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    if (steps < 0) return set_err(AEROS_ERR_ARG, "negative step count");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->step(steps);
}

int32_t aeros_lbm_update_steps(float dt, int32_t force_steps, int32_t steps_per_frame) {
    return update_steps(dt, force_steps, steps_per_frame);
}

int aeros_lbm_update(aeros_lbm* h, float dt, int32_t force_steps, int32_t steps_per_frame) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return run_update(h, update_steps(dt, force_steps, steps_per_frame));
}

int aeros_lbm_lattice_dt(aeros_lbm* h, float* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    *out = lattice_dt_world(h->s);
    return AEROS_OK;
}

int aeros_lbm_frames_start(aeros_lbm* h, uint32_t field_mask) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);   // the worker starts waiting for this lock only after start() returns
    if (h->frames) return set_err(AEROS_ERR_STATE, "frames are already running");
    if (field_mask & ~kFrameFields) return set_err(AEROS_ERR_ARG, "field mask 0x%x names a field not held on the device", field_mask);
    if ((field_mask & ((1u << AEROS_FIELD_T) | (1u << AEROS_FIELD_BUOYANCY))) && !h->s.thermal_ready)
        return set_err(AEROS_ERR_STATE, "thermal module is off");
    (void)hipSetDevice(h->device);
    std::unique_ptr<FrameDriver> d(new (std::nothrow) FrameDriver(h));
    if (!d) return set_err(AEROS_ERR_NOMEM, "frame driver");
    const int rc = d->start(field_mask);
    if (rc) return rc;
    h->frames = std::move(d);
    return AEROS_OK;
}

int aeros_lbm_frame_submit(aeros_lbm* h, float dt, int32_t force_steps, int32_t steps_per_frame) {
    if (!h || !h->frames) return set_err(AEROS_ERR_STATE, "frames are not running");
    return h->frames->submit(FrameDriver::Req{dt, force_steps, steps_per_frame});
}

int aeros_lbm_frame_latest(aeros_lbm* h, aeros_frame* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    if (!h->frames) return set_err(AEROS_ERR_STATE, "frames are not running");
    return h->frames->latest(out);
}

int aeros_lbm_frames_wait(aeros_lbm* h) {
    if (!h || !h->frames) return set_err(AEROS_ERR_STATE, "frames are not running");
    return h->frames->wait();
}

int aeros_lbm_frames_stop(aeros_lbm* h) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    if (!h->frames) return AEROS_OK;
    const int rc = h->frames->wait();
    h->frames.reset();
    return rc;
}

int aeros_lbm_read_field(aeros_lbm* h, int32_t field, float* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->read_field(field, nullptr, out);
}

int aeros_lbm_read_field_f64(aeros_lbm* h, int32_t field, double* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->read_field(field, out, nullptr);
}

int aeros_lbm_read(aeros_lbm* h, uint32_t fields, float* rho, float* ux, float* uy, float* uz, float* T) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    struct { uint32_t bit; int id; float* p; } v[5] = {{AEROS_READ_RHO, AEROS_FIELD_RHO, rho}, {AEROS_READ_UX, AEROS_FIELD_UX, ux},
                                                     {AEROS_READ_UY, AEROS_FIELD_UY, uy}, {AEROS_READ_UZ, AEROS_FIELD_UZ, uz},
                                                     {AEROS_READ_T, AEROS_FIELD_T, T}};
    for (auto& e : v) {
        if (!(fields & e.bit) || !e.p) continue;
        const int rc = h->eng->read_field(e.id, nullptr, e.p);
        if (rc) return rc;
    }
    return AEROS_OK;
}

int aeros_lbm_get_scalars(aeros_lbm* h, aeros_lbm_scalars* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    h->eng->scalars(out);
    return AEROS_OK;
}

int aeros_lbm_read_populations(aeros_lbm* h, double* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->read_populations(out);
}

int aeros_lbm_cv_force(aeros_lbm* h, float hx, float hy, float hz, double out[3]) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->cv_force(hx, hy, hz, out);
}

int aeros_lbm_projected_area(aeros_lbm* h, int32_t axis, float* out) {
    if (!h || !out) return set_err(AEROS_ERR_ARG, "null argument");
    if (axis < 0 || axis > 2) return set_err(AEROS_ERR_ARG, "axis %d", axis);
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->projected_area(axis, out);
}

int aeros_lbm_checksums(aeros_lbm* h, double* rho_sum, double* u_sum) {
    if (!h || !rho_sum || !u_sum) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->checksums(rho_sum, u_sum);
}

int aeros_lbm_rho_sum(aeros_lbm* h, double* rho_sum) {
    double u_sum = 0.0;
    return aeros_lbm_checksums(h, rho_sum, &u_sum);
}

int aeros_lbm_thermal_upload(aeros_lbm* h, const float* T) {
    if (!h || !T) return set_err(AEROS_ERR_ARG, "null argument");
    AEROS_LOCK_HANDLE(h);
    (void)hipSetDevice(h->device);
    return h->eng->thermal_upload(T);
}

int aeros_lbm_thermal_reset(aeros_lbm* h) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);
    if (!h->s.thermal_ready) return set_err(AEROS_ERR_STATE, "thermal module not initialised");
    std::vector<float> T;
    thermal_initial_field(h->s, T);
    (void)hipSetDevice(h->device);
    return h->eng->thermal_upload(T.data());
}

int aeros_lbm_equilibrium_moments(float rho, float ux, float uy, float uz, float* sum_feq, float* mom_x, float* mom_y,
                                  float* mom_z) {
    if (!sum_feq || !mom_x || !mom_y || !mom_z) return 0;
    if (check_device(0)) return 0;
    DevBuf d;
    if (d.alloc(64)) return 0;
    float* o = d.as<float>();
    int* ok = reinterpret_cast<int*>(d.as<char>() + 32);
    hipLaunchKernelGGL(k_eq_moments, dim3(1), dim3(1), 0, 0, rho, ux, uy, uz, o, ok);
    if (hipGetLastError() != hipSuccess) return 0;
    char hb[64];
    if (hipMemcpy(hb, d.p, 64, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    int okv;
    std::memcpy(&okv, hb + 32, sizeof okv);
    if (!okv) return 0;
    float v[4];
    std::memcpy(v, hb, sizeof v);
    *sum_feq = v[0]; *mom_x = v[1]; *mom_y = v[2]; *mom_z = v[3];
    return 1;
}

int aeros_voxelize(const float* tris, int64_t ntris, int32_t nx, int32_t ny, int32_t nz, const float box_min[3],
                   const float cell[3], int32_t* out) {
    if ((!tris && ntris > 0) || !box_min || !cell || !out || nx <= 0 || ny <= 0 || nz <= 0)
        return set_err(AEROS_ERR_ARG, "bad voxelize arguments");
    return voxelize_host(tris, ntris, nx, ny, nz, box_min, cell, out);
}

int aeros_voxelize_gpu(const float* tris, int64_t ntris, int32_t nx, int32_t ny, int32_t nz, const float box_min[3],
                       const float cell[3], int32_t device, int32_t* out) {
    if ((!tris && ntris > 0) || !box_min || !cell || !out || nx <= 0 || ny <= 0 || nz <= 0)
        return set_err(AEROS_ERR_ARG, "bad voxelize arguments");
    int rc = check_device(device);
    if (rc) return rc;
    const int64_t n = (int64_t)nx * ny * nz;
    std::vector<VoxNode> nodes;
    std::vector<VoxTri> vt;
    voxel_bvh(tris, ntris, nodes, vt);
    if (nodes.empty()) {   // no usable triangle: nothing is solid
        std::fill(out, out + n, 0);
        return AEROS_OK;
    }
    float jit[3];
    voxel_jitter(cell, jit);
    DevBuf dn, dt, dout;
    if ((rc = dn.alloc(nodes.size() * sizeof(VoxNode))) || (rc = dt.alloc(vt.size() * sizeof(VoxTri))) ||
        (rc = dout.alloc((size_t)n * sizeof(int32_t))))
        return rc;
    HIP_TRY(hipMemcpy(dn.p, nodes.data(), nodes.size() * sizeof(VoxNode), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dt.p, vt.data(), vt.size() * sizeof(VoxTri), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_voxelize, dim3(grid_for(n)), dim3(256), 0, 0, dn.as<VoxNode>(), dt.as<VoxTri>(), nx, ny, n,
                       box_min[0], box_min[1], box_min[2], cell[0], cell[1], cell[2], jit[0], jit[1], jit[2],
                       dout.as<int32_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return AEROS_OK;
}

int aeros_lbm_vorticity_q(aeros_lbm* h) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    return h->eng->vorticity_q();
}

int aeros_lbm_sample(aeros_lbm* h, int32_t field, const float* pos, int64_t npos, float* out, uint8_t* valid) {
    if (!h || npos < 0 || (npos > 0 && (!pos || !out))) return set_err(AEROS_ERR_ARG, "bad sample arguments");
    AEROS_LOCK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    return h->eng->sample(field, pos, npos, out, valid);
}

int aeros_lbm_sample_noalloc(aeros_lbm* h, int32_t field, const float* pos, int64_t npos, float* out, uint8_t* valid) {
    if (!h || npos < 0 || (npos > 0 && (!pos || !out))) return set_err(AEROS_ERR_ARG, "bad sample arguments");
    if (npos > kNoAllocPoints)
        return set_err(AEROS_ERR_TOO_LARGE, "%lld points exceed the %lld-point staging of the allocation-free sampler",
                       (long long)npos, (long long)kNoAllocPoints);
    AEROS_LOCK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    unsigned char* stage = nullptr;
    if (int rc = noalloc_stage(h->device, &stage)) return rc;
    // handles on one device share its staging: their calls take turns
    std::lock_guard<std::mutex> turn(g_noalloc_use[h->device]);
    return h->eng->sample_staged(field, pos, npos, out, valid, stage);
}

int aeros_lbm_sample_velocity(aeros_lbm* h, int32_t world, const float* pos, int64_t npos, float* out) {
    if (!h || npos < 0 || (npos > 0 && (!pos || !out))) return set_err(AEROS_ERR_ARG, "bad sample arguments");
    AEROS_LOCK_HANDLE(h);
    HIP_TRY(hipSetDevice(h->device));
    return h->eng->sample_velocity(world, pos, npos, out);
}

int aeros_stl_detect(const uint8_t* data, int64_t size, aeros_stl_info* out) {
    if (!out || size < 0 || (!data && size > 0)) return set_err(AEROS_ERR_ARG, "bad STL detect arguments");
    stl_detect(data, (size_t)size, out);
    return AEROS_OK;
}

namespace {
int stl_hand_over(const std::vector<float>& v, const std::vector<float>& nv, float** tris, float** normals,
                  int64_t* ntris) {
    float* a = (float*)std::malloc(std::max<size_t>(v.size(), 1) * sizeof(float));
    float* b = (float*)std::malloc(std::max<size_t>(nv.size(), 1) * sizeof(float));
    if (!a || !b) {
        std::free(a);
        std::free(b);
        return set_err(AEROS_ERR_NOMEM, "STL output allocation failed");
    }
    std::memcpy(a, v.data(), v.size() * sizeof(float));
    std::memcpy(b, nv.data(), nv.size() * sizeof(float));
    *tris = a;
    *normals = b;
    *ntris = (int64_t)(v.size() / 9);
    return AEROS_OK;
}
}  // namespace

int aeros_stl_parse(const uint8_t* data, int64_t size, float** tris, float** normals, int64_t* ntris,
                    int64_t* dropped) {
    if (!tris || !normals || !ntris || size < 0 || (!data && size > 0)) return set_err(AEROS_ERR_ARG, "bad STL arguments");
    *tris = *normals = nullptr;
    *ntris = 0;
    std::vector<float> v, nv;
    int64_t bad = 0;
    const int rc = stl_parse(data, (size_t)size, v, nv, &bad);
    if (dropped) *dropped = bad;
    if (rc) {
        aeros_stl_info info;
        stl_detect(data, (size_t)size, &info);
        return set_err(rc, "STL not loaded: format %d (reason %d), %lld triangles dropped", info.format, info.reason,
                       (long long)bad);
    }
    return stl_hand_over(v, nv, tris, normals, ntris);
}

int aeros_stl_load(const char* path, float** tris, float** normals, int64_t* ntris, int64_t* dropped) {
    if (!path || !tris || !normals || !ntris) return set_err(AEROS_ERR_ARG, "bad STL arguments");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return set_err(AEROS_ERR_ARG, "cannot open %s", path);
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz < 10 || sz > (2L << 30)) {
        std::fclose(fp);
        return set_err(AEROS_ERR_ARG, "STL file size %ld outside 10 B .. 2 GiB", sz);
    }
    std::vector<uint8_t> buf((size_t)sz);
    const size_t got = std::fread(buf.data(), 1, buf.size(), fp);
    std::fclose(fp);
    if (got != buf.size()) return set_err(AEROS_ERR_ARG, "short read of %s (%zu of %ld bytes)", path, got, sz);
    return aeros_stl_parse(buf.data(), (int64_t)buf.size(), tris, normals, ntris, dropped);
}

void aeros_stl_free(float* p) { std::free(p); }

int aeros_stl_write_binary(const char* path, const float* tris, int64_t ntris, const char* header) {
    if (!path || ntris < 0 || ntris > 0xffffffffLL || (!tris && ntris > 0)) return set_err(AEROS_ERR_ARG, "bad STL writer arguments");
    std::vector<uint8_t> buf(84 + (size_t)ntris * 50, 0);
    if (header) std::memcpy(buf.data(), header, std::min<size_t>(std::strlen(header), 80));
    const uint32_t nt = (uint32_t)ntris;
    for (int k = 0; k < 4; ++k) buf[80 + k] = (uint8_t)(nt >> (8 * k));
    for (int64_t t = 0; t < ntris; ++t) {
        const float* a = tris + t * 9;
        const float e1[3] = {a[3] - a[0], a[4] - a[1], a[5] - a[2]};
        const float e2[3] = {a[6] - a[0], a[7] - a[1], a[8] - a[2]};
        float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len > 0.0f && std::isfinite(len)) { n[0] /= len; n[1] /= len; n[2] /= len; }
        else { n[0] = 0; n[1] = 0; n[2] = 1; }
        uint8_t* p = buf.data() + 84 + (size_t)t * 50;
        std::memcpy(p, n, 12);
        std::memcpy(p + 12, a, 36);
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return set_err(AEROS_ERR_ARG, "cannot create %s", path);
    const size_t put = std::fwrite(buf.data(), 1, buf.size(), fp);
    const int closed = std::fclose(fp);
    if (put != buf.size() || closed != 0) return set_err(AEROS_ERR_ARG, "short write of %s", path);
    return AEROS_OK;
}

// buildVoxelGrid's grid (aeros/src/voxel_grid.cpp:250-300, voxelPlan at 172-223)
int aeros_voxel_plan(const float bb_min[3], const float bb_max[3], float max_dim, int32_t res, int32_t dims[3],
                     float box_min[3], float box_max[3]) {
    if (!bb_min || !bb_max || !dims || !box_min || !box_max) return set_err(AEROS_ERR_ARG, "null argument");
    if (res < 8) res = 8;
    if (res > 256) res = 256;
    float margin = 0.14f * max_dim;
    if (!std::isfinite(margin) || margin < 0.001f) margin = 0.1f;
    float size[3];
    for (int i = 0; i < 3; ++i) {
        box_min[i] = bb_min[i] - margin;
        box_max[i] = bb_max[i] + margin;
        size[i] = box_max[i] - box_min[i];
        if (size[i] < 1e-6f) size[i] = 1.0f;
        if (!(size[i] > 1e-6f) || !std::isfinite(size[i])) size[i] = 1.0f;
    }
    const float m = std::max(size[0], std::max(size[1], size[2]));
    dims[0] = res;
    dims[1] = std::max(4, (int)std::floor(res * size[1] / m));
    dims[2] = std::max(4, (int)std::floor(res * size[2] / m));
    return AEROS_OK;
}

int aeros_lbm_last_error(char* buf, int32_t size) {
    const int code = t_err_code;
    if (buf && size > 0) std::snprintf(buf, (size_t)size, "%s", code ? t_err_msg : "");
    t_err_code = 0;
    t_err_msg[0] = 0;
    return code;
}

int aeros_probe(aeros_backend_info* out) {
    if (!out) return set_err(AEROS_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    out->compiled_with_hip = 1;
    int rv = 0, dv = 0;
    (void)hipRuntimeGetVersion(&rv);
    (void)hipDriverGetVersion(&dv);
    out->runtime_version = rv;
    out->driver_version = dv;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) { (void)hipGetLastError(); count = 0; }
    out->device_count = count;
    if (count <= 0) {
        std::snprintf(out->status, sizeof out->status, "no HIP device");
        out->last_error = AEROS_ERR_NO_DEVICE;
        return AEROS_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) { out->last_error = AEROS_ERR_HIP; return AEROS_ERR_HIP; }
    std::snprintf(out->device_name, sizeof out->device_name, "%s", prop.name);
    std::snprintf(out->arch_name, sizeof out->arch_name, "%s", prop.gcnArchName);
    out->gfx_major = prop.major;
    out->gfx_minor = prop.minor;
    out->total_memory_mb = (int64_t)(prop.totalGlobalMem >> 20);
    out->compute_units = prop.multiProcessorCount;
    (void)hipSetDevice(0);
    int* d = nullptr;
    int hv[4] = {1, 2, 3, 4};
    bool ok = hipMalloc(&d, sizeof hv) == hipSuccess;
    ok = ok && hipMemcpy(d, hv, sizeof hv, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_probe, dim3(1), dim3(4), 0, 0, d);
        ok = hipGetLastError() == hipSuccess && hipMemcpy(hv, d, sizeof hv, hipMemcpyDeviceToHost) == hipSuccess;
        ok = ok && hv[0] == 2 && hv[3] == 5;
    }
    if (d) (void)hipFree(d);
    out->usable = ok ? 1 : 0;
    out->last_error = ok ? 0 : AEROS_ERR_HIP;
    std::snprintf(out->status, sizeof out->status, "%s %s, %d CUs, %lld MB: test kernel %s", prop.name,
                  prop.gcnArchName, prop.multiProcessorCount, (long long)out->total_memory_mb, ok ? "ran" : "failed");
    return ok ? AEROS_OK : AEROS_ERR_HIP;
}

int aeros_lbm_kernel_timing(aeros_lbm* h, int32_t enable, double* total_ms, int64_t* launches) {
    if (!h) return set_err(AEROS_ERR_ARG, "null handle");
    AEROS_LOCK_HANDLE(h);
    EngineBase* e = h->eng.get();
    if (enable < 0) { e->timed_ms = 0.0; e->timed_launches = 0; e->timing = 0; }
    else e->timing = enable > 2 ? 2 : enable;
    if (total_ms) *total_ms = e->timed_ms;
    if (launches) *launches = e->timed_launches;
    return AEROS_OK;
}

void aeros_release_all(void) {
    // handles own their buffers; nothing library-wide is cached
}

void setVoxelData(const int* voxel, const float* dist, int nx, int ny, int nz, float mnX, float mnY, float mnZ,
                  float csX, float csY, float csZ) {
    // The reference uploads only the signed distance field here, for its
    // potential-flow colouring (aeros/src/kernel.cu:570-594); the LBM takes
    // its mask through aeros_lbm_create / aeros_lbm_set_mask instead.
    (void)voxel; (void)dist; (void)nx; (void)ny; (void)nz;
    (void)mnX; (void)mnY; (void)mnZ; (void)csX; (void)csY; (void)csZ;
}

}  // extern "C"

#include "lbm_serve.inc"
