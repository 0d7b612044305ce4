// gltf_common.h - pieces the glTF front end shares with the scene generators
// and the glTF exporter, so that a generated world written as .glb and read
// back gives the same arrays bit for bit: the 8192^2 atlas layers and their
// packer, and the normalisation applied to normals and tangents on load.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace ptcore {

// atlas layer side (world.rs:29)
constexpr uint32_t ATLAS = 8192;

// Guillotine packer of one atlas layer: best short-side fit among free
// rectangles, then a split into the strip to the right and the rest below.
struct AtlasRect {
    uint32_t x, y, w, h;
};
struct AtlasLayer {
    std::vector<AtlasRect> free{{0, 0, ATLAS, ATLAS}};
    bool place(uint32_t w, uint32_t h, uint32_t& x, uint32_t& y) {
        size_t best = free.size();
        uint32_t best_fit = std::numeric_limits<uint32_t>::max();
        for (size_t i = 0; i < free.size(); ++i)
            if (free[i].w >= w && free[i].h >= h) {
                uint32_t fit = std::min(free[i].w - w, free[i].h - h);
                if (fit < best_fit) {
                    best_fit = fit;
                    best = i;
                }
            }
        if (best == free.size()) return false;
        AtlasRect r = free[best];
        free.erase(free.begin() + (long)best);
        x = r.x;
        y = r.y;
        if (r.w > w) free.push_back({r.x + w, r.y, r.w - w, h});
        if (r.h > h) free.push_back({r.x, r.y + h, r.w, r.h - h});
        return true;
    }
};

// the normalisation of loaded normals and tangents (glam's Vec3::normalize,
// x * (1 / length) with the length summed x, y, then z)
inline void normalize3(float v[3]) {
    const float inv = 1.0f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] *= inv;
}

// Searches the vectors within 3 ulps per component of `v`, by growing
// Chebyshev shells (v itself first) in a fixed order, for one that `accept`
// (which may adjust it in place) takes; writes it to `out`.
template <class Accept>
bool search_near_ulps(const float v[3], Accept accept, float out[3]) {
    for (int r = 0; r <= 3; ++r)
        for (int dx = -r; dx <= r; ++dx)
            for (int dy = -r; dy <= r; ++dy)
                for (int dz = -r; dz <= r; ++dz) {
                    if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) != r) continue;
                    float c[3] = {v[0], v[1], v[2]};
                    const int d[3] = {dx, dy, dz};
                    for (int k = 0; k < 3; ++k)
                        for (int s = 0; s < std::abs(d[k]); ++s) c[k] = std::nextafter(c[k], d[k] > 0 ? 2.0f : -2.0f);
                    if (accept(c)) {
                        std::memcpy(out, c, sizeof(c));
                        return true;
                    }
                }
    return false;
}

// What loading a stored normal or tangent direction through an identity node
// transform gives: the transform's sums turn -0 into +0, then normalize3.
inline void load_unit(const float in[3], float out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = in[k] + 0.0f;
    normalize3(out);
}

// A unit vector near `v` that loading leaves unchanged: -0 components become
// +0, and the result is a fixed point of normalize3. Repeated normalisation
// ends in a 2-cycle for about one float32 unit vector in six, so the nearest
// fixed point within a few ulps is taken there. Returns false if none is found.
inline bool canonical_unit(float v[3]) {
    float u[3];
    load_unit(v, u);
    for (float& c : u) c = c + 0.0f;
    return search_near_ulps(
        u,
        [](float* c) {
            for (int k = 0; k < 3; ++k) c[k] = c[k] + 0.0f;
            float b[3];
            load_unit(c, b);
            return std::memcmp(c, b, sizeof(b)) == 0;
        },
        v);
}

// A vector that loads back (load_unit) as exactly `n`: n itself when it is a
// fixed point, else the first one found within 3 ulps of n, else within 1 ulp
// of s * n for s = 1 + j/256, j = 1..255, else s * n for s = 1 + j/4096
// (normalisation is close to scale invariant, so a unit vector's preimages
// lie along its ray, at any length).
// Returns false if none is found.
inline bool unit_preimage(const float n[3], float out[3]) {
    auto loads_as_n = [n](float* c) {
        float b[3];
        load_unit(c, b);
        return std::memcmp(n, b, sizeof(b)) == 0;
    };
    if (search_near_ulps(n, loads_as_n, out)) return true;
    for (int j = 1; j < 256; ++j) {
        const float s = 1.0f + (float)j / 256.0f;
        const float base[3] = {s * n[0], s * n[1], s * n[2]};
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    float c[3] = {base[0], base[1], base[2]};
                    const int d[3] = {dx, dy, dz};
                    for (int k = 0; k < 3; ++k)
                        if (d[k]) c[k] = std::nextafter(c[k], d[k] > 0 ? 4.0f : -4.0f);
                    if (loads_as_n(c)) {
                        std::memcpy(out, c, sizeof(c));
                        return true;
                    }
                }
    }
    for (int j = 1; j < 4096; ++j) {
        const float s = 1.0f + (float)j / 4096.0f;
        float c[3] = {s * n[0], s * n[1], s * n[2]};
        if (loads_as_n(c)) {
            std::memcpy(out, c, sizeof(c));
            return true;
        }
    }
    return false;
}

}  // namespace ptcore
