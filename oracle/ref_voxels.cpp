// TEST INFRASTRUCTURE ONLY.  The reference's voxel vote (its BVH and its
// three-ray majority test, aeros/src/mesh_bvh.h, used as
// aeros/src/voxel_grid.cpp:313-351 uses them) over a caller-given grid, so
// that tests/golden/make_golden.py can record the reference's own mask of
// the synthetic ellipsoid STLs.  Built by oracle/ref.mk into oracle/_ref/.
//
//   oracle/_ref/ref_voxels <tris.f32> <nx> <ny> <nz> <minx> <miny> <minz> <csx> <csy> <csz> <mask.u8>
//
// tris.f32 holds 9 float32 per triangle; the mask is nx*ny*nz bytes, x fastest.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mesh_bvh.h"

int main(int argc, char** argv) {
    if (argc != 12) {
        std::fprintf(stderr, "usage: %s tris.f32 nx ny nz minx miny minz csx csy csz mask.u8\n", argv[0]);
        return 2;
    }
    std::vector<float> v;
    {
        FILE* f = std::fopen(argv[1], "rb");
        if (!f) return 1;
        float buf[4096];
        size_t k;
        while ((k = std::fread(buf, sizeof(float), 4096, f)) > 0) v.insert(v.end(), buf, buf + k);
        std::fclose(f);
    }
    const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
    const float mn[3] = {(float)std::atof(argv[5]), (float)std::atof(argv[6]), (float)std::atof(argv[7])};
    const float cs[3] = {(float)std::atof(argv[8]), (float)std::atof(argv[9]), (float)std::atof(argv[10])};

    // the triangle filter of the reference's voxelizer: finite, non-degenerate
    std::vector<meshbvh::Triangle> tris;
    for (size_t i = 0; i + 8 < v.size(); i += 9) {
        bool ok = true;
        for (int j = 0; j < 9; ++j) ok = ok && std::isfinite(v[i + j]);
        if (!ok) continue;
        const meshbvh::Triangle t = meshbvh::makeTriangle(glm::vec3(v[i], v[i + 1], v[i + 2]),
                                                          glm::vec3(v[i + 3], v[i + 4], v[i + 5]),
                                                          glm::vec3(v[i + 6], v[i + 7], v[i + 8]));
        if (!meshbvh::isDegenerate(t)) tris.push_back(t);
    }
    meshbvh::BVH bvh;
    if (!tris.empty()) bvh.build(tris);
    const glm::vec3 cell(cs[0], cs[1], cs[2]);
    std::vector<unsigned char> mask((size_t)nx * ny * nz, 0);
    if (!tris.empty()) {
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const glm::vec3 p(mn[0] + (i + 0.5f) * cs[0], mn[1] + (j + 0.5f) * cs[1], mn[2] + (k + 0.5f) * cs[2]);
                    if (bvh.inside(p, cell)) mask[((size_t)k * ny + j) * nx + i] = 1;
                }
    }
    FILE* f = std::fopen(argv[11], "wb");
    if (!f) return 1;
    const bool ok = std::fwrite(mask.data(), 1, mask.size(), f) == mask.size();
    std::fclose(f);
    size_t solid = 0;
    for (unsigned char m : mask) solid += m;
    std::printf("triangles %zu solid %zu\n", tris.size(), solid);
    return ok ? 0 : 1;
}
