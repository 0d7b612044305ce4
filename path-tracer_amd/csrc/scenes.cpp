// scenes.cpp - seeded procedural scenes for the benchmark workloads.
//
// Output is a triangle soup in "input order" plus materials, atlases and an
// un-normalised sky, i.e. what the reference's World::new sees after glTF
// flattening (src/render/world.rs:1114-1337). All randomness comes from a
// splitmix64 stream of the seed, so a (kind, seed, size) triple always gives
// the same arrays.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/pt_scenes.h"
#include "gltf_common.h"
#include "pt_internal.h"

namespace {

struct Rng64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }  // [0, 1)
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    uint32_t below(uint32_t n) { return (uint32_t)(next() % n); }
};

struct V {
    float x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dotv(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V crossv(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V norm(V a) {
    float l = std::sqrt(dotv(a, a));
    return l > 0.0f ? a * (1.0f / l) : V{0.0f, 1.0f, 0.0f};
}

PtMaterial default_material() {
    // GpuMaterial defaults (world.rs:199-309) with metalness 0 and diffuse roughness 0.5
    PtMaterial m;
    std::memset(&m, 0, sizeof(m));
    for (int k = 0; k < 4; ++k) m.base_color[k] = 1.0f;
    m.base_weight = 1.0f;
    m.base_metalness = 0.0f;
    m.base_diffuse_roughness = 0.5f;
    m.emission_weight = 1.0f;
    m.emission_color_and_luminance[3] = 1.0f;
    for (int k = 0; k < 4; ++k) m.specular_color_and_weight[k] = 1.0f;
    m.specular_roughness = 0.3f;
    m.specular_haze_spread = 0.3f;
    m.specular_ior = 1.5f;
    m.normal_scale = 1.0f;
    m.alpha_mode = PT_ALPHA_MODE_OPAQUE;
    m.alpha_cutoff = 0.5f;
    m.base_color_tex_layer = m.normal_tex_layer = m.metallic_roughness_tex_layer = -1;
    m.emission_tex_layer = m.specular_color_tex_layer = m.specular_weight_tex_layer = -1;
    return m;
}

// a tangent perpendicular to n
V any_tangent(V n) {
    V t = std::fabs(n.x) < 0.9f ? V{1.0f, 0.0f, 0.0f} : V{0.0f, 0.0f, 1.0f};
    return norm(t - n * dotv(n, t));
}

}  // namespace

struct pt_generated {
    std::vector<PtTriangleGeometry> geo;
    std::vector<PtTriangleAttribute> attr;
    std::vector<PtMaterial> mats;
    std::vector<uint8_t> srgb, linear;
    uint32_t srgb_count = 0, linear_count = 0, atlas_size = 0;
    std::vector<float> sky;
    uint32_t sky_w = 0, sky_h = 0;
    float sky_luminance = 1.0f;
    float cam_pos[3] = {NAN, NAN, NAN};
    float yaw = NAN, pitch = NAN, fov = NAN, aspect = 1.0f;
    float tex_uv[4][4] = {};  // atlas offset and scale of the procedural texture pairs

    void tri(V a, V b, V c, V na, V nb, V nc, V ta, V tb, V tc, float uv[6], uint32_t mat) {
        PtTriangleGeometry g = {{a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z}};
        PtTriangleAttribute at;
        std::memset(&at, 0, sizeof(at));
        at.material = mat;
        V ns[3] = {na, nb, nc}, ts[3] = {ta, tb, tc};
        for (int k = 0; k < 3; ++k) {
            at.n[3 * k] = ns[k].x;
            at.n[3 * k + 1] = ns[k].y;
            at.n[3 * k + 2] = ns[k].z;
        }
        std::memcpy(at.uv, uv, sizeof(at.uv));
        float* t[3] = {at.t0, at.t1, at.t2};
        for (int k = 0; k < 3; ++k) {
            t[k][0] = ts[k].x;
            t[k][1] = ts[k].y;
            t[k][2] = ts[k].z;
            t[k][3] = 1.0f;
        }
        geo.push_back(g);
        attr.push_back(at);
    }
    void flat_tri(V a, V b, V c, uint32_t mat) {
        V n = norm(crossv(b - a, c - a));
        V t = any_tangent(n);
        float uv[6] = {0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 1.0f};
        tri(a, b, c, n, n, n, t, t, t, uv, mat);
    }
    void quad(V a, V b, V c, V d, uint32_t mat) {  // a b c d counter-clockwise
        flat_tri(a, b, c, mat);
        flat_tri(a, c, d, mat);
    }
    void const_sky(uint32_t w, uint32_t h, float grey) {
        sky_w = w;
        sky_h = h;
        sky.assign((size_t)w * h * 4, grey);
        for (size_t i = 3; i < sky.size(); i += 4) sky[i] = 1.0f;
    }
};

namespace {

// ------------------------------------------------------------------ noise
struct ValueNoise {
    std::vector<float> lattice;
    uint32_t n = 256;
    explicit ValueNoise(Rng64& r) : lattice(256 * 256) {
        for (auto& v : lattice) v = r.uniform(-1.0f, 1.0f);
    }
    float at(int x, int y) const { return lattice[((uint32_t)y & 255u) * 256u + ((uint32_t)x & 255u)]; }
    static float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
    float sample(float x, float y) const {
        float fx = std::floor(x), fy = std::floor(y);
        int ix = (int)fx, iy = (int)fy;
        float tx = smooth(x - fx), ty = smooth(y - fy);
        float a = at(ix, iy), b = at(ix + 1, iy), c = at(ix, iy + 1), d = at(ix + 1, iy + 1);
        return (a + (b - a) * tx) + ((c + (d - c) * tx) - (a + (b - a) * tx)) * ty;
    }
    float fbm(float x, float y, int octaves) const {
        float sum = 0.0f, amp = 0.5f, freq = 1.0f;
        for (int o = 0; o < octaves; ++o) {
            sum += amp * sample(x * freq, y * freq);
            amp *= 0.5f;
            freq *= 2.0f;
        }
        return sum;
    }
};

// Perlin's improved gradient noise (2002) on a seeded permutation
struct Perlin {
    int p[512];
    explicit Perlin(Rng64& r) {
        int perm[256];
        for (int i = 0; i < 256; ++i) perm[i] = i;
        for (int i = 255; i > 0; --i) std::swap(perm[i], perm[r.below((uint32_t)i + 1)]);
        for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
    }
    static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
    static float grad(int h, float x, float y) {
        switch (h & 7) {
            case 0: return x + y;
            case 1: return -x + y;
            case 2: return x - y;
            case 3: return -x - y;
            case 4: return x;
            case 5: return -x;
            case 6: return y;
            default: return -y;
        }
    }
    float noise(float x, float y) const {
        float fx = std::floor(x), fy = std::floor(y);
        int X = (int)fx & 255, Y = (int)fy & 255;
        x -= fx;
        y -= fy;
        float u = fade(x), v = fade(y);
        int aa = p[p[X] + Y], ab = p[p[X] + Y + 1], ba = p[p[X + 1] + Y], bb = p[p[X + 1] + Y + 1];
        float l0 = grad(aa, x, y) + u * (grad(ba, x - 1.0f, y) - grad(aa, x, y));
        float l1 = grad(ab, x, y - 1.0f) + u * (grad(bb, x - 1.0f, y - 1.0f) - grad(ab, x, y - 1.0f));
        return l0 + v * (l1 - l0);
    }
    // periodic fBm over a tile of `period` lattice cells so textures wrap
    float fbm(float x, float y, int octaves) const {
        float s = 0.0f, a = 0.5f, f = 1.0f;
        for (int o = 0; o < octaves; ++o) {
            s += a * noise(x * f, y * f);
            a *= 0.5f;
            f *= 2.0f;
        }
        return s;
    }
};

// ------------------------------------------------------------------ builders
void heightfield(pt_generated& g, uint32_t grid, float extent, float amp, const ValueNoise& noise, float noise_scale,
                 uint32_t mat, float uv_tiles) {
    std::vector<V> pos((size_t)grid * grid), nrm((size_t)grid * grid);
    auto h = [&](float x, float z) { return amp * noise.fbm(x * noise_scale + 17.3f, z * noise_scale + 5.1f, 5); };
    const float step = 2.0f * extent / (float)(grid - 1);
    for (uint32_t j = 0; j < grid; ++j)
        for (uint32_t i = 0; i < grid; ++i) {
            float x = -extent + step * (float)i, z = -extent + step * (float)j;
            pos[(size_t)j * grid + i] = {x, h(x, z), z};
        }
    for (uint32_t j = 0; j < grid; ++j)
        for (uint32_t i = 0; i < grid; ++i) {
            uint32_t i0 = i ? i - 1 : i, i1 = i + 1 < grid ? i + 1 : i;
            uint32_t j0 = j ? j - 1 : j, j1 = j + 1 < grid ? j + 1 : j;
            V dx = pos[(size_t)j * grid + i1] - pos[(size_t)j * grid + i0];
            V dz = pos[(size_t)j1 * grid + i] - pos[(size_t)j0 * grid + i];
            nrm[(size_t)j * grid + i] = norm(crossv(dz, dx));
        }
    auto uvof = [&](uint32_t i, uint32_t j, float* out) {
        out[0] = (float)i / (float)(grid - 1) * uv_tiles;
        out[1] = (float)j / (float)(grid - 1) * uv_tiles;
    };
    for (uint32_t j = 0; j + 1 < grid; ++j)
        for (uint32_t i = 0; i + 1 < grid; ++i) {
            size_t a = (size_t)j * grid + i, b = a + 1, c = a + grid + 1, d = a + grid;
            uint32_t ij[4][2] = {{i, j}, {i + 1, j}, {i + 1, j + 1}, {i, j + 1}};
            size_t idx[4] = {a, b, c, d};
            V tan[4];
            for (int k = 0; k < 4; ++k) tan[k] = any_tangent(nrm[idx[k]]);
            float uv[6];
            int t0[3] = {0, 2, 1}, t1[3] = {0, 3, 2};
            for (int* t : {t0, t1}) {
                for (int k = 0; k < 3; ++k) uvof(ij[t[k]][0], ij[t[k]][1], uv + 2 * k);
                g.tri(pos[idx[t[0]]], pos[idx[t[1]]], pos[idx[t[2]]], nrm[idx[t[0]]], nrm[idx[t[1]]],
                      nrm[idx[t[2]]], tan[t[0]], tan[t[1]], tan[t[2]], uv, mat);
            }
        }
}

struct Icosphere {
    std::vector<V> verts;
    std::vector<uint32_t> faces;
    explicit Icosphere(int level) {
        const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
        V base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
        for (auto& v : base) verts.push_back(norm(v));
        faces = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
        for (int l = 0; l < level; ++l) {
            std::vector<uint32_t> nf;
            std::vector<std::pair<uint64_t, uint32_t>> cache;
            auto mid = [&](uint32_t a, uint32_t b) {
                uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
                for (auto& e : cache)
                    if (e.first == key) return e.second;
                verts.push_back(norm((verts[a] + verts[b]) * 0.5f));
                cache.push_back({key, (uint32_t)verts.size() - 1});
                return (uint32_t)verts.size() - 1;
            };
            for (size_t f = 0; f < faces.size(); f += 3) {
                uint32_t a = faces[f], b = faces[f + 1], c = faces[f + 2];
                uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
                uint32_t nfc[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
                nf.insert(nf.end(), nfc, nfc + 12);
            }
            faces.swap(nf);
        }
    }
};

void add_sphere(pt_generated& g, const Icosphere& ico, V center, float radius, float rot[9], uint32_t mat) {
    auto xf = [&](V v) {
        return V{rot[0] * v.x + rot[1] * v.y + rot[2] * v.z, rot[3] * v.x + rot[4] * v.y + rot[5] * v.z,
                 rot[6] * v.x + rot[7] * v.y + rot[8] * v.z};
    };
    for (size_t f = 0; f < ico.faces.size(); f += 3) {
        V n[3], p[3], t[3];
        float uv[6];
        for (int k = 0; k < 3; ++k) {
            V local = ico.verts[ico.faces[f + k]];
            n[k] = xf(local);
            p[k] = center + n[k] * radius;
            t[k] = any_tangent(n[k]);
            uv[2 * k] = 0.5f + std::atan2(local.z, local.x) * 0.15915494f;
            uv[2 * k + 1] = std::acos(std::fmax(-1.0f, std::fmin(1.0f, local.y))) * 0.31830989f;
        }
        g.tri(p[0], p[1], p[2], n[0], n[1], n[2], t[0], t[1], t[2], uv, mat);
    }
}

void random_rotation(Rng64& r, float m[9]) {
    float a = r.uniform(0.0f, 6.2831853f), b = r.uniform(0.0f, 6.2831853f), c = r.uniform(-1.0f, 1.0f);
    float sa = std::sin(a), ca = std::cos(a), sb = std::sin(b), cb = std::cos(b), sc = std::sqrt(1.0f - c * c);
    // rotate by a about y, then tilt the pole to (sc cos b, c, sc sin b)
    float ry[9] = {ca, 0, sa, 0, 1, 0, -sa, 0, ca};
    V up = {sc * cb, c, sc * sb};
    V ax = norm(crossv(V{0, 1, 0}, up));
    if (dotv(ax, ax) < 0.5f) ax = {1, 0, 0};
    V bz = crossv(ax, up);
    float tilt[9] = {ax.x, up.x, bz.x, ax.y, up.y, bz.y, ax.z, up.z, bz.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            m[i * 3 + j] = tilt[i * 3 + 0] * ry[0 * 3 + j] + tilt[i * 3 + 1] * ry[1 * 3 + j] + tilt[i * 3 + 2] * ry[2 * 3 + j];
}

float lum(const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

// Give `count` random triangles out of [first, end) their own emissive
// material with flux log-uniform over [flux_lo, flux_hi]:
// strength = flux / (2 pi area luminance(color)) (world.rs:1697-1711).
void make_emitters(pt_generated& g, Rng64& r, size_t first, size_t end, uint32_t count, float flux_lo, float flux_hi) {
    std::vector<uint32_t> chosen;
    const size_t n = end - first;
    if (n == 0) return;
    count = (uint32_t)std::min<size_t>(count, n);
    std::vector<uint8_t> used(n, 0);
    while (chosen.size() < count) {
        uint32_t i = (uint32_t)(r.next() % n);
        if (used[i]) continue;
        used[i] = 1;
        chosen.push_back((uint32_t)(first + i));
    }
    for (uint32_t ti : chosen) {
        PtMaterial m = g.mats[g.attr[ti].material];
        float color[3] = {1.0f, r.uniform(0.75f, 1.0f), r.uniform(0.55f, 0.95f)};
        const float* v = g.geo[ti].v;
        V e1 = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2 = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float area = 0.5f * std::sqrt(dotv(crossv(e1, e2), crossv(e1, e2)));
        float flux = flux_lo * std::exp(r.uniform() * std::log(flux_hi / flux_lo));
        m.emission_color_and_luminance[0] = color[0];
        m.emission_color_and_luminance[1] = color[1];
        m.emission_color_and_luminance[2] = color[2];
        m.emission_color_and_luminance[3] = flux / (6.2831853f * std::fmax(area, 1e-12f) * lum(color));
        m.emission_weight = 1.0f;
        g.attr[ti].material = (uint32_t)g.mats.size();
        g.mats.push_back(m);
    }
}

void gradient_sky(pt_generated& g, uint32_t w, uint32_t h, bool sun, bool rgbe) {
    g.sky_w = w;
    g.sky_h = h;
    g.sky.assign((size_t)w * h * 4, 1.0f);
    for (uint32_t y = 0; y < h; ++y) {
        float theta = ((float)y + 0.5f) / (float)h * 3.14159265f;  // 0 at the zenith
        float up = std::cos(theta);
        for (uint32_t x = 0; x < w; ++x) {
            float phi = ((float)x + 0.5f) / (float)w * 6.2831853f - 3.14159265f;
            float c[3];
            if (up >= 0.0f) {
                float t = std::pow(up, 0.5f);
                c[0] = 0.85f + (0.25f - 0.85f) * t;
                c[1] = 0.9f + (0.45f - 0.9f) * t;
                c[2] = 1.0f;
            } else {
                c[0] = 0.25f;
                c[1] = 0.22f;
                c[2] = 0.2f;
            }
            if (sun) {
                // sun at elevation 35 degrees, azimuth 0.6 rad
                float st = std::sin(theta);
                V d = {st * std::cos(phi), up, st * std::sin(phi)};
                V s = norm(V{std::cos(0.6109f) * std::cos(0.6f), std::sin(0.6109f), std::cos(0.6109f) * std::sin(0.6f)});
                float cosang = dotv(d, s);
                if (cosang > 0.9995f) {
                    c[0] += 400.0f;
                    c[1] += 380.0f;
                    c[2] += 340.0f;
                } else if (cosang > 0.99f) {
                    float k = (cosang - 0.99f) / 0.0095f;
                    c[0] += 6.0f * k;
                    c[1] += 5.5f * k;
                    c[2] += 5.0f * k;
                }
            }
            float* p = &g.sky[((size_t)y * w + x) * 4];
            if (rgbe) {
                // quantise to what an RGBE pixel decodes to (mantissa byte q, exponent byte
                // e + 128: q * 2^(e - 8), as the image crate's Rgbe8Pixel::to_hdr gives it),
                // so the sky can be written as a Radiance .hdr without loss
                float mx = std::fmax(c[0], std::fmax(c[1], c[2]));
                if (mx < 1e-32f) {
                    p[0] = p[1] = p[2] = 0.0f;
                } else {
                    int e;
                    std::frexp(mx, &e);
                    for (int k = 0; k < 3; ++k) {
                        int q = (int)std::ldexp(c[k], 8 - e);
                        p[k] = std::ldexp((float)q, e - 8);
                    }
                }
            } else {
                p[0] = c[0];
                p[1] = c[1];
                p[2] = c[2];
            }
            p[3] = 1.0f;
        }
    }
}

void cornell(pt_generated& g, Rng64& r) {
    // 7 diffuse materials (5 walls, 2 boxes) + 1 emitter; pure EON (specular weight 0)
    for (int i = 0; i < 7; ++i) {
        PtMaterial m = default_material();
        for (int k = 0; k < 3; ++k) m.base_color[k] = r.uniform(0.1f, 0.8f);
        m.base_diffuse_roughness = r.uniform(0.0f, 1.0f);
        m.specular_color_and_weight[3] = 0.0f;
        g.mats.push_back(m);
    }
    PtMaterial light = default_material();
    for (int k = 0; k < 3; ++k) light.base_color[k] = 0.8f;
    light.specular_color_and_weight[3] = 0.0f;
    light.emission_color_and_luminance[0] = light.emission_color_and_luminance[1] = light.emission_color_and_luminance[2] = 1.0f;
    light.emission_color_and_luminance[3] = 15.0f;
    g.mats.push_back(light);
    // walls of [-1, 1]^3, open at +z
    g.quad({-1, -1, 1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, -1}, 0);  // floor
    g.quad({-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}, 1);      // ceiling
    g.quad({-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, 2);  // back
    g.quad({-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, 3);  // left
    g.quad({1, -1, -1}, {1, -1, 1}, {1, 1, 1}, {1, 1, -1}, 4);      // right
    auto box = [&](V c, V half, float angle_deg, uint32_t mat) {
        float a = angle_deg * 3.14159265f / 180.0f, ca = std::cos(a), sa = std::sin(a);
        auto P = [&](float sx, float sy, float sz) {
            float x = sx * half.x, z = sz * half.z;
            return V{c.x + ca * x + sa * z, c.y + sy * half.y, c.z - sa * x + ca * z};
        };
        V p[8] = {P(-1, -1, -1), P(1, -1, -1), P(1, -1, 1), P(-1, -1, 1),
                  P(-1, 1, -1), P(1, 1, -1), P(1, 1, 1), P(-1, 1, 1)};
        g.quad(p[0], p[3], p[2], p[1], mat);  // bottom
        g.quad(p[4], p[5], p[6], p[7], mat);  // top
        g.quad(p[0], p[1], p[5], p[4], mat);  // back
        g.quad(p[3], p[7], p[6], p[2], mat);  // front
        g.quad(p[0], p[4], p[7], p[3], mat);  // left
        g.quad(p[1], p[2], p[6], p[5], mat);  // right
    };
    box({-0.35f, -0.4f, -0.3f}, {0.3f, 0.6f, 0.3f}, 17.0f, 5);
    box({0.35f, -0.7f, 0.3f}, {0.3f, 0.3f, 0.3f}, -17.0f, 6);
    // ceiling light, 2 triangles just below the ceiling
    g.quad({-0.3f, 0.99f, -0.3f}, {0.3f, 0.99f, -0.3f}, {0.3f, 0.99f, 0.3f}, {-0.3f, 0.99f, 0.3f}, 7);
    g.const_sky(64, 32, 0.5f);
    g.sky_luminance = 0.5f;
    g.cam_pos[0] = 0.0f;
    g.cam_pos[1] = 0.0f;
    g.cam_pos[2] = 3.5f;
    g.yaw = 0.0f;
    g.pitch = 0.0f;
    g.fov = 45.0f;
}

// heightfield + icospheres scene with `spheres` spheres of subdivision `level`
void field_scene(pt_generated& g, Rng64& r, uint32_t grid, uint32_t spheres, int level, float extent, float amp,
                 uint32_t emitters, float flux_lo, float flux_hi, int textured_pairs) {
    ValueNoise noise(r);
    PtMaterial ground = default_material();
    ground.base_color[0] = 0.45f;
    ground.base_color[1] = 0.42f;
    ground.base_color[2] = 0.38f;
    ground.specular_roughness = 0.6f;
    ground.base_diffuse_roughness = 0.8f;
    if (textured_pairs > 0) {
        ground.base_color_tex_layer = 0;
        ground.metallic_roughness_tex_layer = 0;
        std::memcpy(ground.base_color_uv, g.tex_uv[0], sizeof(ground.base_color_uv));
        std::memcpy(ground.metallic_roughness_uv, g.tex_uv[0], sizeof(ground.metallic_roughness_uv));
    }
    g.mats.push_back(ground);
    const uint32_t sphere_mats = std::min<uint32_t>(64, std::max<uint32_t>(1, spheres));
    for (uint32_t i = 0; i < sphere_mats; ++i) {
        PtMaterial m = default_material();
        for (int k = 0; k < 3; ++k) m.base_color[k] = r.uniform(0.1f, 0.9f);
        m.specular_roughness = r.uniform(0.05f, 0.9f);
        m.base_metalness = r.uniform() < 0.3f ? 1.0f : 0.0f;
        m.base_diffuse_roughness = r.uniform(0.0f, 1.0f);
        if (textured_pairs > 1 && i % 4 != 3) {
            int pair = 1 + (int)(i % 3);
            m.base_color_tex_layer = 0;
            m.metallic_roughness_tex_layer = 0;
            std::memcpy(m.base_color_uv, g.tex_uv[pair], sizeof(m.base_color_uv));
            std::memcpy(m.metallic_roughness_uv, g.tex_uv[pair], sizeof(m.metallic_roughness_uv));
        }
        g.mats.push_back(m);
    }
    heightfield(g, grid, extent, amp, noise, 1.5f / extent * 2.0f, 0, textured_pairs > 0 ? 8.0f : 1.0f);
    const size_t sphere_first = g.geo.size();
    Icosphere ico(level);
    for (uint32_t s = 0; s < spheres; ++s) {
        float x = r.uniform(-extent * 0.9f, extent * 0.9f), z = r.uniform(-extent * 0.9f, extent * 0.9f);
        float radius = r.uniform(0.02f, 0.06f) * extent;
        float y = amp * noise.fbm(x * (1.5f / extent * 2.0f) + 17.3f, z * (1.5f / extent * 2.0f) + 5.1f, 5) + 0.6f * radius;
        float rot[9];
        random_rotation(r, rot);
        add_sphere(g, ico, {x, y, z}, radius, rot, 1 + s % sphere_mats);
    }
    make_emitters(g, r, sphere_first < g.geo.size() ? sphere_first : 0, g.geo.size(), emitters, flux_lo, flux_hi);
}

void perlin_atlases(pt_generated& g, Rng64& r, uint32_t tex_size) {
    // one sRGB atlas layer with 4 albedo textures and one linear layer with 4
    // metallic-roughness textures, each tex_size^2, placed where the glTF
    // loader's packer puts 4 such images in that order; the rest of each
    // layer is zero, as the loader leaves it
    const uint32_t A = ptcore::ATLAS;
    g.atlas_size = A;
    g.srgb.assign((size_t)A * A * 4, 0);
    g.linear.assign((size_t)A * A * 4, 0);
    g.srgb_count = g.linear_count = 1;
    ptcore::AtlasLayer packer;
    for (int t = 0; t < 4; ++t) {
        Perlin pa(r), pr(r);
        float base[3] = {r.uniform(0.2f, 0.9f), r.uniform(0.2f, 0.9f), r.uniform(0.2f, 0.9f)};
        float tint[3] = {r.uniform(0.1f, 0.9f), r.uniform(0.1f, 0.9f), r.uniform(0.1f, 0.9f)};
        uint32_t ox = 0, oy = 0;
        packer.place(tex_size, tex_size, ox, oy);
        g.tex_uv[t][0] = (float)ox / (float)A;
        g.tex_uv[t][1] = (float)oy / (float)A;
        g.tex_uv[t][2] = g.tex_uv[t][3] = (float)tex_size / (float)A;
        const float freq = 8.0f / (float)tex_size;
        for (uint32_t y = 0; y < tex_size; ++y)
            for (uint32_t x = 0; x < tex_size; ++x) {
                float n = 0.5f + 0.7f * pa.fbm((float)x * freq, (float)y * freq, 5);
                float m = 0.5f + 0.7f * pr.fbm((float)x * freq * 0.5f + 31.0f, (float)y * freq * 0.5f, 4);
                n = std::fmin(std::fmax(n, 0.0f), 1.0f);
                m = std::fmin(std::fmax(m, 0.0f), 1.0f);
                uint8_t* s = &g.srgb[(((size_t)(oy + y) * A) + ox + x) * 4];
                uint8_t* l = &g.linear[(((size_t)(oy + y) * A) + ox + x) * 4];
                for (int k = 0; k < 3; ++k) s[k] = (uint8_t)(255.0f * (base[k] + (tint[k] - base[k]) * n) + 0.5f);
                s[3] = 255;
                l[0] = 0;
                l[1] = (uint8_t)(255.0f * (0.15f + 0.85f * m) + 0.5f);  // G: roughness factor
                l[2] = (uint8_t)(m > 0.8f ? 255 : 0);                    // B: metalness factor
                l[3] = 255;
            }
    }
}

// Make the world one that a glTF file can hold exactly: written as .glb with
// explicit normals and tangents and read back by the scene loader, it gives
// the same arrays bit for bit.
// - materials are numbered in order of first use by the triangles, as the
//   loader creates them (unused ones are dropped);
// - no coordinate is -0 (the loader's identity transform turns it into +0),
//   and normals and tangents are fixed points of the loader's normalisation;
// - a triangle with the tangent (1, 0, 0) and w = 1 at a corner, the value the
//   loader takes for "no tangents" and replaces by MikkTSpace tangents, gets
//   the same frame written as tangent (-1, 0, 0) with w = -1 at every corner
//   (bitangent = cross(n, t) * w is unchanged).
bool canonicalize(pt_generated& g) {
    std::vector<int64_t> remap(g.mats.size(), -1);
    std::vector<PtMaterial> mats;
    for (PtTriangleAttribute& at : g.attr) {
        if (remap[at.material] < 0) {
            remap[at.material] = (int64_t)mats.size();
            mats.push_back(g.mats[at.material]);
        }
        at.material = (uint32_t)remap[at.material];
    }
    g.mats.swap(mats);
    for (PtTriangleGeometry& tri : g.geo)
        for (float& c : tri.v) c = c + 0.0f;
    for (PtTriangleAttribute& at : g.attr) {
        float* t[3] = {at.t0, at.t1, at.t2};
        for (int k = 0; k < 3; ++k)
            if (!ptcore::canonical_unit(&at.n[3 * k]) || !ptcore::canonical_unit(t[k])) return false;
    }
    return true;
}

}  // namespace

extern "C" {

pt_status pt_generate_scene(int kind, uint64_t seed, uint32_t triangles, uint32_t width, uint32_t height,
                            pt_generated** out) {
    using ptcore::set_error;
    if (!out) return set_error(PT_ERR_INVALID, "pt_generate_scene: null output");
    *out = nullptr;
    if (width == 0 || height == 0) return set_error(PT_ERR_INVALID, "pt_generate_scene: zero image size");
    pt_generated* g = new (std::nothrow) pt_generated();
    if (!g) return set_error(PT_ERR_MEMORY, "out of host memory");
    Rng64 r{seed * 0x2545F4914F6CDD1Dull + (uint64_t)kind};
    try {
        switch (kind) {
            case PT_SCENE_CORNELL:
                cornell(*g, r);
                break;
            case PT_SCENE_FIELD:
                field_scene(*g, r, 301, 64, 3, 1.0f, 0.2f, 256, 1.0f, 1000.0f, 0);
                gradient_sky(*g, 128, 64, false, false);
                g->sky_luminance = 200.0f;
                g->cam_pos[0] = 0.0f;
                g->cam_pos[1] = 0.9f;
                g->cam_pos[2] = 1.9f;
                g->yaw = 0.0f;
                g->pitch = -25.0f;
                g->fov = 55.0f;
                break;
            case PT_SCENE_LARGE:
                perlin_atlases(*g, r, 2048);
                field_scene(*g, r, 1202, 4096, 2, 8.0f, 0.8f, 4096, 1.0f, 1000.0f, 4);
                gradient_sky(*g, 1024, 512, true, true);
                g->sky_luminance = 500.0f;
                g->cam_pos[0] = 0.0f;
                g->cam_pos[1] = 2.5f;
                g->cam_pos[2] = 9.0f;
                g->yaw = 0.0f;
                g->pitch = -15.0f;
                g->fov = 60.0f;
                break;
            case PT_SCENE_SIZED: {
                if (triangles < 64) return delete g, set_error(PT_ERR_INVALID, "PT_SCENE_SIZED needs >= 64 triangles");
                uint32_t spheres = (uint32_t)std::max<double>(1.0, std::floor(0.3 * triangles / 320.0));
                uint32_t rest = triangles > spheres * 320 ? triangles - spheres * 320 : 2;
                uint32_t grid = (uint32_t)std::floor(std::sqrt((double)rest / 2.0)) + 1;
                grid = std::max<uint32_t>(grid, 2);
                uint32_t emitters = std::min<uint32_t>(256, std::max<uint32_t>(16, triangles / 2000));
                field_scene(*g, r, grid, spheres, 2, 1.0f, 0.2f, emitters, 1.0f, 1000.0f, 0);
                gradient_sky(*g, 128, 64, false, false);
                g->sky_luminance = 200.0f;
                g->cam_pos[0] = 0.0f;
                g->cam_pos[1] = 0.9f;
                g->cam_pos[2] = 1.9f;
                g->yaw = 0.0f;
                g->pitch = -25.0f;
                g->fov = 55.0f;
                break;
            }
            default:
                delete g;
                return set_error(PT_ERR_INVALID, "unknown scene kind %d", kind);
        }
    } catch (const std::bad_alloc&) {
        delete g;
        return set_error(PT_ERR_MEMORY, "out of host memory while generating a scene");
    }
    if (!canonicalize(*g)) {
        delete g;
        return set_error(PT_ERR_STATE, "no load-stable normal near a generated one");
    }
    g->aspect = (float)width / (float)height;
    *out = g;
    return PT_OK;
}

pt_status pt_generated_world_desc(const pt_generated* g, pt_world_desc* d) {
    if (!g || !d) return ptcore::set_error(PT_ERR_INVALID, "pt_generated_world_desc: null argument");
    std::memset(d, 0, sizeof(*d));
    d->geometries = g->geo.data();
    d->attributes = g->attr.data();
    d->triangle_count = (uint32_t)g->geo.size();
    d->materials = g->mats.data();
    d->material_count = (uint32_t)g->mats.size();
    d->srgb_atlases = g->srgb.empty() ? nullptr : g->srgb.data();
    d->srgb_atlas_count = g->srgb_count;
    d->linear_atlases = g->linear.empty() ? nullptr : g->linear.data();
    d->linear_atlas_count = g->linear_count;
    d->atlas_size = g->atlas_size;
    d->skybox_rgba = g->sky.data();
    d->skybox_width = g->sky_w;
    d->skybox_height = g->sky_h;
    d->skybox_luminance = g->sky_luminance;
    for (int k = 0; k < 3; ++k) d->camera_position[k] = g->cam_pos[k];
    d->camera_yaw_degrees = g->yaw;
    d->camera_pitch_degrees = g->pitch;
    d->camera_fov_degrees = g->fov;
    d->aspect_ratio = g->aspect;
    return PT_OK;
}

pt_status pt_generated_destroy(pt_generated* g) {
    delete g;
    return PT_OK;
}

}  // extern "C"
