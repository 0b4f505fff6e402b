// STL input and output for the set-mask producer (host code).
//
// Format detection restates stlformat::detect (aeros/src/stl_format.h:86-144):
// ASCII when the file starts with "solid" and its first 4 KiB of text hold a
// facet/vertex keyword; otherwise binary when the 32-bit count at byte 80
// promises triangles that the file holds (a trailing tail is allowed and
// counted); a truncated binary file and text without keywords are rejected.
// Parsing restates loadSTL (aeros/src/stl_loader.cpp:30-205): binary normals
// are sanitised and normalised, triangles with a non-finite coordinate are
// dropped; ASCII vertices are read in threes, non-finite ones dropped, and the
// normal of each triangle is recomputed from its edges.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_common.h"

namespace aeros {

namespace {

uint32_t le32(const unsigned char* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

bool starts_with_solid(const unsigned char* d, size_t n) {
    if (n < 5) return false;
    static const char s[] = "solid";
    for (int i = 0; i < 5; ++i) {
        char c = (char)d[i];
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
        if (c != s[i]) return false;
    }
    return true;
}

// no NUL and no control byte other than \t..\r in the first 2 KiB
bool text_like(const unsigned char* d, size_t n) {
    const size_t lim = n < 2048 ? n : 2048;
    for (size_t i = 0; i < lim; ++i) {
        const unsigned char c = d[i];
        if (c == 0 || c < 9 || (c > 13 && c < 32) || c == 127) return false;
    }
    return true;
}

bool has_keywords(const unsigned char* d, size_t n) {
    const size_t lim = n < 4096 ? n : 4096;
    auto lc = [&](size_t i) { return (unsigned char)(d[i] | 0x20); };
    for (size_t i = 0; i + 5 < lim; ++i) {
        if (lc(i) == 'f' && lc(i + 1) == 'a' && lc(i + 2) == 'c' && lc(i + 3) == 'e' && lc(i + 4) == 't') return true;
        if (lc(i) == 'v' && lc(i + 1) == 'e' && lc(i + 2) == 'r' && lc(i + 3) == 't' && lc(i + 4) == 'e' &&
            lc(i + 5) == 'x')
            return true;
    }
    return false;
}

void normalise(float n[3], float fx, float fy, float fz, float eps) {
    const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > eps) { n[0] /= len; n[1] /= len; n[2] /= len; }
    else { n[0] = fx; n[1] = fy; n[2] = fz; }
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

}  // namespace

void stl_detect(const unsigned char* d, size_t size, aeros_stl_info* out) {
    std::memset(out, 0, sizeof *out);
    out->format = AEROS_STL_UNKNOWN;
    if (!d || size == 0) { out->reason = AEROS_STL_EMPTY; return; }
    const bool solid = starts_with_solid(d, size);
    const bool text = text_like(d, size);
    const bool keywords = text && has_keywords(d, size);
    const uint32_t nt = size >= 84 ? le32(d + 80) : 0;
    const uint64_t need = 84ull + (uint64_t)nt * 50ull;
    const bool fits = nt > 0 && need <= (uint64_t)size;
    out->declared_triangles = nt;
    if (solid && keywords) { out->format = AEROS_STL_ASCII; out->reason = AEROS_STL_SOLID_KEYWORDS; return; }
    if (fits) {
        out->format = AEROS_STL_BINARY;
        out->payload_bytes = (int64_t)nt * 50;
        out->trailing_bytes = (int64_t)(size - need);
        out->reason = out->trailing_bytes ? AEROS_STL_BINARY_TAIL : AEROS_STL_BINARY_EXACT;
        return;
    }
    if (text && !solid) {
        out->format = keywords ? AEROS_STL_ASCII : AEROS_STL_UNKNOWN;
        out->reason = keywords ? AEROS_STL_TEXT_NO_SOLID : AEROS_STL_TEXT_NO_KEYWORDS;
        return;
    }
    if (nt > 0 && need > (uint64_t)size) { out->reason = AEROS_STL_TRUNCATED; return; }
    out->reason = AEROS_STL_UNRECOGNISED;
}

int stl_parse(const unsigned char* d, size_t size, std::vector<float>& verts, std::vector<float>& normals,
              int64_t* dropped) {
    aeros_stl_info info;
    stl_detect(d, size, &info);
    verts.clear();
    normals.clear();
    int64_t bad = 0;
    if (info.format == AEROS_STL_BINARY) {
        const uint32_t nt = info.declared_triangles;
        verts.reserve((size_t)nt * 9);
        normals.reserve((size_t)nt * 9);
        size_t off = 84;
        for (uint32_t i = 0; i < nt; ++i, off += 50) {
            float n[3], v[9];
            std::memcpy(n, d + off, 12);
            if (!finite3(n)) { n[0] = 0; n[1] = 0; n[2] = 1; }
            normalise(n, 0, 0, 1, 1e-6f);
            std::memcpy(v, d + off + 12, 36);
            if (!finite3(v) || !finite3(v + 3) || !finite3(v + 6)) { ++bad; continue; }
            for (int k = 0; k < 3; ++k) {
                verts.insert(verts.end(), v + 3 * k, v + 3 * k + 3);
                normals.insert(normals.end(), n, n + 3);
            }
        }
    } else if (info.format == AEROS_STL_ASCII) {
        std::vector<float> tv;
        std::string line;
        size_t pos = 0;
        while (pos < size) {
            size_t e = pos;
            while (e < size && d[e] != '\n') ++e;
            line.assign((const char*)d + pos, e - pos);
            pos = e + 1;
            size_t i = 0;
            while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
            size_t j = i;
            while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
            std::string tok = line.substr(i, j - i);
            for (char& c : tok) c = (char)std::tolower((unsigned char)c);
            if (tok != "vertex") continue;
            const char* p = line.c_str() + j;
            float v[3] = {0, 0, 0};
            bool ok = true;
            for (int k = 0; k < 3 && ok; ++k) {
                char* endp = nullptr;
                v[k] = std::strtof(p, &endp);
                ok = endp != p;
                p = endp;
            }
            if (!ok || !finite3(v)) { ++bad; continue; }
            tv.insert(tv.end(), v, v + 3);
        }
        const size_t tris = tv.size() / 9;
        verts.reserve(tris * 9);
        normals.reserve(tris * 9);
        for (size_t t = 0; t < tris; ++t) {
            const float* a = &tv[t * 9];
            const float e1[3] = {a[3] - a[0], a[4] - a[1], a[5] - a[2]};
            const float e2[3] = {a[6] - a[0], a[7] - a[1], a[8] - a[2]};
            float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            normalise(n, 0, 1, 0, 1e-9f);
            for (int k = 0; k < 3; ++k) {
                verts.insert(verts.end(), a + 3 * k, a + 3 * k + 3);
                normals.insert(normals.end(), n, n + 3);
            }
        }
    } else {
        return AEROS_ERR_ARG;
    }
    if (dropped) *dropped = bad;
    return verts.empty() ? AEROS_ERR_ARG : AEROS_OK;
}

}  // namespace aeros
