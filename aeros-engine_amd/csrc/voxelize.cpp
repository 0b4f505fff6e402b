// Set-mask producer: solid voxels by the three-ray majority vote against a
// triangle BVH, at the centres of an arbitrary regular grid.  Restates
// meshbvh (aeros/src/mesh_bvh.h:56-270) and the voxel loop of buildVoxelGrid
// (aeros/src/voxel_grid.cpp:313-351): median split on the longest centroid
// axis, leaves of at most 4 triangles, depth <= 48, float slab tests,
// double-precision Moller-Trumbore with a relative determinant epsilon, and
// rays jittered by (0.731, 0.377, 0.557) * 1e-3 of a cell.  Runs on the host
// threads at setup time (once per body).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

#include "lbm_common.h"

namespace aeros {

namespace {

struct Vec { float v[3]; };

struct Tri {
    Vec a, b, c;
    Vec lo, hi, mid;
};

struct Node {
    Vec lo, hi;
    int left = -1, right = -1, first = 0, count = 0;
};

class TriBvh {
public:
    explicit TriBvh(std::vector<Tri> t) : tris_(std::move(t)) {
        if (!tris_.empty()) build(0, (int)tris_.size(), 0);
    }
    bool empty() const { return nodes_.empty(); }

    void flatten(std::vector<VoxNode>& nodes, std::vector<VoxTri>& tris) const {
        nodes.resize(nodes_.size());
        for (size_t i = 0; i < nodes_.size(); ++i) {
            const Node& a = nodes_[i];
            VoxNode& b = nodes[i];
            for (int k = 0; k < 3; ++k) { b.lo[k] = a.lo.v[k]; b.hi[k] = a.hi.v[k]; }
            b.left = a.left; b.right = a.right; b.first = a.first; b.count = a.count;
        }
        tris.resize(tris_.size());
        for (size_t i = 0; i < tris_.size(); ++i)
            for (int k = 0; k < 3; ++k) {
                tris[i].a[k] = tris_[i].a.v[k]; tris[i].b[k] = tris_[i].b.v[k]; tris[i].c[k] = tris_[i].c.v[k];
            }
    }

    // parity of the number of crossings of the ray o + t d, t > 1e-7
    int crossings(const Vec& o, const Vec& d) const {
        const float inf = std::numeric_limits<float>::infinity();
        float inv[3];
        for (int k = 0; k < 3; ++k) inv[k] = d.v[k] != 0.0f ? 1.0f / d.v[k] : inf;
        if (!box_hit(o, inv, nodes_[0])) return 0;
        int stack[128];
        int sp = 0, hits = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node& nd = nodes_[stack[--sp]];
            if (nd.count) {
                for (int i = nd.first; i < nd.first + nd.count; ++i) hits += hit(o, d, tris_[i]);
            } else {
                if (box_hit(o, inv, nodes_[nd.left])) stack[sp++] = nd.left;
                if (box_hit(o, inv, nodes_[nd.right])) stack[sp++] = nd.right;
            }
        }
        return hits;
    }

private:
    std::vector<Tri> tris_;
    std::vector<Node> nodes_;

    int build(int first, int count, int depth) {
        const int id = (int)nodes_.size();
        nodes_.emplace_back();
        Vec lo = tris_[first].lo, hi = tris_[first].hi, cl = tris_[first].mid, ch = tris_[first].mid;
        for (int i = first + 1; i < first + count; ++i)
            for (int k = 0; k < 3; ++k) {
                lo.v[k] = std::min(lo.v[k], tris_[i].lo.v[k]);
                hi.v[k] = std::max(hi.v[k], tris_[i].hi.v[k]);
                cl.v[k] = std::min(cl.v[k], tris_[i].mid.v[k]);
                ch.v[k] = std::max(ch.v[k], tris_[i].mid.v[k]);
            }
        nodes_[id].lo = lo;
        nodes_[id].hi = hi;
        int axis = 0;
        const float e0 = ch.v[0] - cl.v[0], e1 = ch.v[1] - cl.v[1], e2 = ch.v[2] - cl.v[2];
        if (e1 > e0) axis = 1;
        if (e2 > (axis ? e1 : e0)) axis = 2;
        const float ext = axis == 0 ? e0 : (axis == 1 ? e1 : e2);
        if (count <= 4 || depth >= 48 || !(ext > 1e-12f)) {
            nodes_[id].first = first;
            nodes_[id].count = count;
            return id;
        }
        const int half = count / 2;
        std::nth_element(tris_.begin() + first, tris_.begin() + first + half, tris_.begin() + first + count,
                         [axis](const Tri& p, const Tri& q) { return p.mid.v[axis] < q.mid.v[axis]; });
        const int l = build(first, half, depth + 1);
        const int r = build(first + half, count - half, depth + 1);
        nodes_[id].left = l;
        nodes_[id].right = r;
        return id;
    }

    static bool box_hit(const Vec& o, const float inv[3], const Node& nd) {
        float t0 = -std::numeric_limits<float>::infinity(), t1 = std::numeric_limits<float>::infinity();
        for (int k = 0; k < 3; ++k) {
            float a = (nd.lo.v[k] - o.v[k]) * inv[k], b = (nd.hi.v[k] - o.v[k]) * inv[k];
            if (a > b) std::swap(a, b);
            if (std::isnan(a) || std::isnan(b)) continue;   // ray in the slab plane
            t0 = std::max(t0, a);
            t1 = std::min(t1, b);
            if (t1 < t0) return false;
        }
        return t1 >= 0.0f;
    }

    static int hit(const Vec& o, const Vec& d, const Tri& t) {
        const double e1[3] = {(double)t.b.v[0] - t.a.v[0], (double)t.b.v[1] - t.a.v[1], (double)t.b.v[2] - t.a.v[2]};
        const double e2[3] = {(double)t.c.v[0] - t.a.v[0], (double)t.c.v[1] - t.a.v[1], (double)t.c.v[2] - t.a.v[2]};
        const double dd[3] = {d.v[0], d.v[1], d.v[2]};
        const double p[3] = {dd[1] * e2[2] - dd[2] * e2[1], dd[2] * e2[0] - dd[0] * e2[2], dd[0] * e2[1] - dd[1] * e2[0]};
        const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
        const double scale = std::fabs(e1[0] * e2[1] - e1[1] * e2[0]) + std::fabs(e1[1] * e2[2] - e1[2] * e2[1]) +
                             std::fabs(e1[2] * e2[0] - e1[0] * e2[2]) + 1e-30;
        if (std::fabs(det) < 1e-12 * scale) return 0;
        const double inv = 1.0 / det;
        const double s[3] = {(double)o.v[0] - t.a.v[0], (double)o.v[1] - t.a.v[1], (double)o.v[2] - t.a.v[2]};
        const double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
        if (u < 0.0 || u > 1.0) return 0;
        const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
        const double v = (dd[0] * q[0] + dd[1] * q[1] + dd[2] * q[2]) * inv;
        if (v < 0.0 || u + v > 1.0) return 0;
        const double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
        return (tt > 1e-7 && tt < 1e30) ? 1 : 0;
    }
};

}  // namespace

namespace {
std::vector<Tri> usable_triangles(const float* v, int64_t ntris) {
    std::vector<Tri> tris;
    tris.reserve((size_t)std::max<int64_t>(ntris, 0));
    for (int64_t t = 0; t < ntris; ++t) {
        const float* p = v + t * 9;
        bool ok = true;
        for (int k = 0; k < 9; ++k) ok = ok && std::isfinite(p[k]);
        if (!ok) continue;
        Tri tr;
        for (int k = 0; k < 3; ++k) {
            tr.a.v[k] = p[k]; tr.b.v[k] = p[3 + k]; tr.c.v[k] = p[6 + k];
            tr.lo.v[k] = std::min(tr.a.v[k], std::min(tr.b.v[k], tr.c.v[k]));
            tr.hi.v[k] = std::max(tr.a.v[k], std::max(tr.b.v[k], tr.c.v[k]));
            tr.mid.v[k] = (tr.a.v[k] + tr.b.v[k] + tr.c.v[k]) * (1.0f / 3.0f);
        }
        // degenerate test |(b-a) x (c-a)| <= 1e-14 (aeros/src/mesh_bvh.h:51-54)
        const float ux = tr.b.v[0] - tr.a.v[0], uy = tr.b.v[1] - tr.a.v[1], uz = tr.b.v[2] - tr.a.v[2];
        const float wx = tr.c.v[0] - tr.a.v[0], wy = tr.c.v[1] - tr.a.v[1], wz = tr.c.v[2] - tr.a.v[2];
        const float cx = uy * wz - wy * uz, cy = uz * wx - wz * ux, cz = ux * wy - wx * uy;
        const float len = std::sqrt(cx * cx + cy * cy + cz * cz);
        if (!(len > 1e-14f) || !std::isfinite(len)) continue;
        tris.push_back(tr);
    }
    return tris;
}
}  // namespace

void voxel_jitter(const float cs[3], float jit[3]) {
    jit[0] = 0.731f * 1e-3f * cs[0];
    jit[1] = 0.377f * 1e-3f * cs[1];
    jit[2] = 0.557f * 1e-3f * cs[2];
}

void voxel_bvh(const float* v, int64_t ntris, std::vector<VoxNode>& nodes, std::vector<VoxTri>& tris) {
    const TriBvh bvh(usable_triangles(v, ntris));
    nodes.clear();
    tris.clear();
    if (!bvh.empty()) bvh.flatten(nodes, tris);
}

int voxelize_host(const float* v, int64_t ntris, int nx, int ny, int nz, const float mn[3], const float cs[3],
                  int32_t* out) {
    const TriBvh bvh(usable_triangles(v, ntris));
    const int64_t total = (int64_t)nx * ny * nz;
    if (bvh.empty()) {
        std::fill(out, out + total, 0);
        return AEROS_OK;
    }
    float jit[3];
    voxel_jitter(cs, jit);
    const float jx = jit[0], jy = jit[1], jz = jit[2];
    const Vec dx{{1, 0, 0}}, dy{{0, 1, 0}}, dz{{0, 0, 1}};
    auto work = [&](int64_t rowStart, int64_t rowEnd) {
        for (int64_t r = rowStart; r < rowEnd; ++r) {   // r = k*ny + j
            const int k = (int)(r / ny), j = (int)(r % ny);
            const float pz = mn[2] + (k + 0.5f) * cs[2], py = mn[1] + (j + 0.5f) * cs[1];
            for (int i = 0; i < nx; ++i) {
                const float px = mn[0] + (i + 0.5f) * cs[0];
                int votes = 0;
                votes += bvh.crossings(Vec{{px, py + jy, pz + jz}}, dx) & 1;
                votes += bvh.crossings(Vec{{px + jz, py, pz + jx}}, dy) & 1;
                votes += bvh.crossings(Vec{{px + jy, py + jx, pz}}, dz) & 1;
                out[r * nx + i] = votes >= 2 ? 1 : 0;
            }
        }
    };
    const int64_t rows = (int64_t)nz * ny;
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (rows < 64) nt = 1;
    std::vector<std::thread> pool;
    const int64_t chunk = (rows + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t a = t * chunk, b = std::min(rows, a + chunk);
        if (a < b) pool.emplace_back(work, a, b);
    }
    for (auto& th : pool) th.join();
    return AEROS_OK;
}

}  // namespace aeros
