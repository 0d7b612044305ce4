// gltf_ingest.cpp - the scene and sky front end: glTF 2.0 (.glb, or .gltf
// with embedded data URIs) to the triangle soup, materials and texture
// atlases pt_world_build takes, and sky image files to RGBA32F.
//
// Restates Scene::new (src/render/world.rs:887-1423) and
// GpuMaterial::new (world.rs:1990-2155):
//   * every image becomes RGBA8 by the reference's format rules (R8 -> rrr1,
//     RG8 -> rg01, RGB8 -> rgb1, 16-bit -> high bytes, world.rs:946-1010) and
//     goes into an sRGB atlas when a material uses it as base colour,
//     emission or specular colour texture, else into a linear atlas
//     (8192^2 layers, world.rs:29, 898-1075);
//   * nodes are flattened depth-first from the default scene's roots (all
//     nodes when there is no default scene) with a LIFO stack; a node's
//     matrix is parent * T * R * S (world.rs:1076-1135);
//   * positions by the model matrix; normals by its inverse transpose,
//     normalised; tangents by the model matrix, normalised, w kept; a
//     tangent of exactly (1,0,0,1) or with w varying inside a triangle, or
//     missing tangents, or missing normals (then flat normals), switch the
//     primitive to MikkTSpace tangents with w negated (world.rs:1137-1266);
//   * texcoords default to the cycle (0,0),(1,0),(0,1) (world.rs:1107-1111);
//   * materials are created at first use and shared per glTF material
//     (world.rs:1268-1288), with OpenPBR extras in [0, 1] (world.rs:2046-2096).
// Matrix arithmetic follows glam's evaluation order (column sums
// ((x + y) + z) + w, no fused multiply-add). Atlas placement is this
// library's own guillotine packer, not guillotiere's: texel contents and the
// sampled values are unchanged, only the atlas offsets differ.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <new>
#include <set>
#include <string>
#include <vector>

#include "../../include/pt_assets.h"
#include "gltf_common.h"
#include "image_io.h"
#include "json_lite.h"
#include "pt_internal.h"
#include "tangents.h"

namespace ptcore {

struct Assets {
    std::vector<PtTriangleGeometry> geometries;
    std::vector<PtTriangleAttribute> attributes;
    std::vector<PtMaterial> materials;
    std::vector<uint8_t> srgb, linear;
    uint32_t srgb_layers = 0, linear_layers = 0;
    FloatImage sky;
    std::vector<std::string> warnings;
};

namespace {

bool fail(std::string& err, const std::string& m) {
    err = m;
    return false;
}

// ------------------------------------------------------------- glam-order matrix maths
struct M4 {
    float c[4][4];  // c[column][row]
};
M4 identity() {
    M4 m{};
    for (int i = 0; i < 4; ++i) m.c[i][i] = 1.0f;
    return m;
}
void mul_vec(const M4& a, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r) out[r] = ((a.c[0][r] * v[0] + a.c[1][r] * v[1]) + a.c[2][r] * v[2]) + a.c[3][r] * v[3];
}
M4 mul(const M4& a, const M4& b) {
    M4 o;
    for (int j = 0; j < 4; ++j) mul_vec(a, b.c[j], o.c[j]);
    return o;
}
M4 from_quat(float x, float y, float z, float w) {
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
    const float wx = w * x2, wy = w * y2, wz = w * z2;
    M4 m = identity();
    const float cols[3][3] = {{1.0f - (yy + zz), xy + wz, xz - wy},
                              {xy - wz, 1.0f - (xx + zz), yz + wx},
                              {xz + wy, yz - wx, 1.0f - (xx + yy)}};
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) m.c[j][r] = cols[j][r];
    return m;
}
// GLM-style cofactor inverse, the formulation glam uses
M4 inverse(const M4& M) {
    auto m = [&](int c, int r) { return M.c[c][r]; };
    const float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
    const float c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3), c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    const float c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3), c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    const float c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
    const float c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2), c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    const float c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3), c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    const float c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
    const float c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2), c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    const float c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1), c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    const float fac0[4] = {c00, c00, c02, c03}, fac1[4] = {c04, c04, c06, c07}, fac2[4] = {c08, c08, c10, c11};
    const float fac3[4] = {c12, c12, c14, c15}, fac4[4] = {c16, c16, c18, c19}, fac5[4] = {c20, c20, c22, c23};
    const float vec0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)}, vec1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
    const float vec2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)}, vec3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
    const float sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
    M4 inv;
    for (int k = 0; k < 4; ++k) {
        inv.c[0][k] = ((vec1[k] * fac0[k] - vec2[k] * fac1[k]) + vec3[k] * fac2[k]) * sa[k];
        inv.c[1][k] = ((vec0[k] * fac0[k] - vec2[k] * fac3[k]) + vec3[k] * fac4[k]) * sb[k];
        inv.c[2][k] = ((vec0[k] * fac1[k] - vec1[k] * fac3[k]) + vec3[k] * fac5[k]) * sa[k];
        inv.c[3][k] = ((vec0[k] * fac2[k] - vec1[k] * fac4[k]) + vec2[k] * fac5[k]) * sb[k];
    }
    const float row0[4] = {inv.c[0][0], inv.c[1][0], inv.c[2][0], inv.c[3][0]};
    const float det = ((M.c[0][0] * row0[0] + M.c[0][1] * row0[1]) + M.c[0][2] * row0[2]) + M.c[0][3] * row0[3];
    const float rcp = 1.0f / det;
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) inv.c[j][k] *= rcp;
    return inv;
}
M4 transpose(const M4& a) {
    M4 t;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) t.c[i][j] = a.c[j][i];
    return t;
}
void point3(const M4& m, const float p[3], float o[3]) {
    for (int r = 0; r < 3; ++r) o[r] = ((m.c[0][r] * p[0] + m.c[1][r] * p[1]) + m.c[2][r] * p[2]) + m.c[3][r];
}
void vector3(const M4& m, const float v[3], float o[3]) {
    for (int r = 0; r < 3; ++r) o[r] = (m.c[0][r] * v[0] + m.c[1][r] * v[1]) + m.c[2][r] * v[2];
}

// ------------------------------------------------------------- buffers and accessors
struct Doc {
    Json j;
    std::vector<std::vector<uint8_t>> buffers;
};

bool base64(const std::string& s, std::vector<uint8_t>& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    uint32_t acc = 0;
    int bits = 0;
    for (char c : s) {
        if (c == '=' || c == '\n' || c == '\r') continue;
        int v = val(c);
        if (v < 0) return false;
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return true;
}

bool data_uri(const std::string& uri, std::vector<uint8_t>& out, std::string& err) {
    size_t comma = uri.find(',');
    if (uri.rfind("data:", 0) != 0 || comma == std::string::npos || uri.find(";base64") > comma)
        return fail(err, "external URI '" + uri.substr(0, 64) + "' is not supported (embed the data or use .glb)");
    if (!base64(uri.substr(comma + 1), out)) return fail(err, "bad base64 data URI");
    return true;
}

struct View {
    const uint8_t* p = nullptr;
    size_t len = 0, stride = 0;
};

bool buffer_view(const Doc& d, long vi, View& v, std::string& err) {
    const Json* views = d.j.get("bufferViews");
    if (!views || views->type != Json::ARR || vi < 0 || (size_t)vi >= views->arr.size())
        return fail(err, "bufferView index out of range");
    const Json& bv = views->arr[vi];
    const long b = bv.index_at("buffer");
    if (b < 0 || (size_t)b >= d.buffers.size()) return fail(err, "buffer index out of range");
    const double off = bv.num_at("byteOffset", 0), len = bv.num_at("byteLength", -1);
    const std::vector<uint8_t>& buf = d.buffers[b];
    if (off < 0 || len < 0 || off + len > (double)buf.size()) return fail(err, "bufferView outside its buffer");
    v.p = buf.data() + (size_t)off;
    v.len = (size_t)len;
    v.stride = (size_t)bv.num_at("byteStride", 0);
    return true;
}

struct Accessor {
    const uint8_t* base = nullptr;
    size_t count = 0, stride = 0;
    int comp = 0, ncomp = 0;
    bool normalized = false;
    std::vector<uint8_t> zeros;  // accessors without a bufferView read as zeros
    double at(size_t i, int c) const {
        const uint8_t* q = base + i * stride;
        switch (comp) {
            case 5120: { int8_t v; std::memcpy(&v, q + c, 1); return normalized ? std::fmax(v / 127.0, -1.0) : v; }
            case 5121: { uint8_t v = q[c]; return normalized ? v / 255.0 : v; }
            case 5122: { int16_t v; std::memcpy(&v, q + 2 * c, 2); return normalized ? std::fmax(v / 32767.0, -1.0) : v; }
            case 5123: { uint16_t v; std::memcpy(&v, q + 2 * c, 2); return normalized ? v / 65535.0 : v; }
            case 5125: { uint32_t v; std::memcpy(&v, q + 4 * c, 4); return v; }
            default: { float v; std::memcpy(&v, q + 4 * c, 4); return v; }
        }
    }
    float f32(size_t i, int c) const {  // as gltf's reader gives floats (into_f32 for normalized integers)
        const uint8_t* q = base + i * stride;
        if (comp == 5126) {
            float v;
            std::memcpy(&v, q + 4 * c, 4);
            return v;
        }
        if (comp == 5121) return (float)q[c] / 255.0f;
        if (comp == 5123) {
            uint16_t v;
            std::memcpy(&v, q + 2 * c, 2);
            return (float)v / 65535.0f;
        }
        return (float)at(i, c);
    }
    uint32_t u32(size_t i) const {
        const uint8_t* q = base + i * stride;
        if (comp == 5121) return q[0];
        if (comp == 5123) {
            uint16_t v;
            std::memcpy(&v, q, 2);
            return v;
        }
        uint32_t v;
        std::memcpy(&v, q, 4);
        return v;
    }
};

bool accessor(const Doc& d, long ai, Accessor& a, std::string& err, int want_ncomp = 0) {
    const Json* accs = d.j.get("accessors");
    if (!accs || accs->type != Json::ARR || ai < 0 || (size_t)ai >= accs->arr.size())
        return fail(err, "accessor index out of range");
    const Json& A = accs->arr[ai];
    if (A.get("sparse")) return fail(err, "sparse accessors are not supported");
    a.comp = (int)A.num_at("componentType", 0);
    int csize;
    switch (a.comp) {
        case 5120: case 5121: csize = 1; break;
        case 5122: case 5123: csize = 2; break;
        case 5125: case 5126: csize = 4; break;
        default: return fail(err, "bad accessor componentType");
    }
    const Json* type = A.get("type");
    const std::string t = type && type->type == Json::STR ? type->str : "";
    a.ncomp = t == "SCALAR" ? 1 : t == "VEC2" ? 2 : t == "VEC3" ? 3 : t == "VEC4" ? 4 : t == "MAT4" ? 16 : 0;
    if (!a.ncomp) return fail(err, "unsupported accessor type '" + t + "'");
    if (want_ncomp && a.ncomp != want_ncomp) return fail(err, "accessor has the wrong type for its attribute");
    const Json* norm = A.get("normalized");
    a.normalized = norm && norm->type == Json::BOOL && norm->b;
    const double cnt = A.num_at("count", -1);
    if (cnt < 0 || cnt > 1e9) return fail(err, "bad accessor count");
    a.count = (size_t)cnt;
    const size_t elem = (size_t)csize * a.ncomp;
    const long vi = A.index_at("bufferView");
    if (vi < 0) {
        a.zeros.assign(elem, 0);
        a.base = a.zeros.data();
        a.stride = 0;
        return true;
    }
    View v;
    if (!buffer_view(d, vi, v, err)) return false;
    const double off = A.num_at("byteOffset", 0);
    a.stride = v.stride ? v.stride : elem;
    if (off < 0 || (a.count && off + (double)a.stride * (a.count - 1) + elem > (double)v.len))
        return fail(err, "accessor outside its bufferView");
    a.base = v.p + (size_t)off;
    return true;
}

// ------------------------------------------------------------- materials
struct ImageSlot {
    int layer = -1;
    float uv[4] = {0, 0, 0, 0};
};

PtMaterial make_material(const Json* mat, const std::map<long, ImageSlot>& srgb, const std::map<long, ImageSlot>& lin,
                         const Doc& d, std::vector<std::string>& warnings) {
    PtMaterial m{};
    auto tex_image = [&](const Json* texinfo) -> long {
        if (!texinfo) return -1;
        long ti = texinfo->index_at("index");
        const Json* texs = d.j.get("textures");
        if (ti < 0 || !texs || (size_t)ti >= texs->arr.size()) return -1;
        return texs->arr[ti].index_at("source");
    };
    auto bind = [](const std::map<long, ImageSlot>& slots, long img, int32_t& layer, float uv[4]) {
        auto it = img >= 0 ? slots.find(img) : slots.end();
        if (it == slots.end()) {
            layer = -1;
            for (int k = 0; k < 4; ++k) uv[k] = 0.0f;
        } else {
            layer = it->second.layer;
            for (int k = 0; k < 4; ++k) uv[k] = it->second.uv[k];
        }
    };
    const Json* pbr = mat ? mat->get("pbrMetallicRoughness") : nullptr;
    const Json* ext = mat ? mat->get("extensions") : nullptr;
    const Json* spec = ext ? ext->get("KHR_materials_specular") : nullptr;
    const Json* bcf = pbr ? pbr->get("baseColorFactor") : nullptr;
    for (int k = 0; k < 4; ++k)
        m.base_color[k] = bcf && bcf->type == Json::ARR && bcf->arr.size() == 4 ? (float)bcf->arr[k].num_or(1.0) : 1.0f;
    m.base_metalness = pbr ? (float)pbr->num_at("metallicFactor", 1.0) : 1.0f;
    m.base_diffuse_roughness = pbr ? (float)pbr->num_at("roughnessFactor", 1.0) : 1.0f;
    bind(srgb, tex_image(pbr ? pbr->get("baseColorTexture") : nullptr), m.base_color_tex_layer, m.base_color_uv);
    bind(lin, tex_image(pbr ? pbr->get("metallicRoughnessTexture") : nullptr), m.metallic_roughness_tex_layer,
         m.metallic_roughness_uv);
    const Json* ntex = mat ? mat->get("normalTexture") : nullptr;
    bind(lin, tex_image(ntex), m.normal_tex_layer, m.normal_uv);
    bind(srgb, tex_image(mat ? mat->get("emissiveTexture") : nullptr), m.emission_tex_layer, m.emission_uv);
    bind(lin, tex_image(spec ? spec->get("specularTexture") : nullptr), m.specular_weight_tex_layer, m.specular_weight_uv);
    bind(srgb, tex_image(spec ? spec->get("specularColorTexture") : nullptr), m.specular_color_tex_layer,
         m.specular_color_uv);
    // OpenPBR parameters from the material's extras (world.rs:2046-2096)
    float base_weight = 1.0f, emission_weight = 1.0f, spec_rough = 0.3f, aniso = 0.0f, haze = 0.0f, haze_spread = 0.3f,
          retro = 0.0f;
    if (const Json* ex = mat ? mat->get("extras") : nullptr) {
        if (ex->type == Json::OBJ) {
            for (const auto& kv : ex->obj) {
                float* target = kv.first == "OpenPBR_base_weight"                     ? &base_weight
                                : kv.first == "OpenPBR_emission_weight"               ? &emission_weight
                                : kv.first == "OpenPBR_specular_roughness"            ? &spec_rough
                                : kv.first == "OpenPBR_specular_roughness_anisotropy" ? &aniso
                                : kv.first == "OpenPBR_specular_haze"                 ? &haze
                                : kv.first == "OpenPBR_specular_haze_spread"          ? &haze_spread
                                : kv.first == "OpenPBR_specular_retroreflectivity"    ? &retro
                                                                                      : nullptr;
                if (!target) {
                    warnings.push_back("unrecognized key in material extras: " + kv.first);
                    continue;
                }
                if (!kv.second.is_num()) {
                    warnings.push_back(kv.first + " value is not a float");
                    continue;
                }
                const float v = (float)kv.second.num;
                if (v >= 0.0f && v <= 1.0f) *target = v;
                else warnings.push_back(kv.first + " value is out of range [0, 1]");
            }
        } else {
            warnings.push_back("material extras is not a JSON object");
        }
    }
    m.base_weight = base_weight;
    m.emission_weight = emission_weight;
    const Json* ef = mat ? mat->get("emissiveFactor") : nullptr;
    for (int k = 0; k < 3; ++k)
        m.emission_color_and_luminance[k] = ef && ef->type == Json::ARR && ef->arr.size() == 3 ? (float)ef->arr[k].num_or(0.0) : 0.0f;
    const Json* es = ext ? ext->get("KHR_materials_emissive_strength") : nullptr;
    m.emission_color_and_luminance[3] = es ? (float)es->num_at("emissiveStrength", 1.0) : 1.0f;
    const Json* scf = spec ? spec->get("specularColorFactor") : nullptr;
    for (int k = 0; k < 3; ++k)
        m.specular_color_and_weight[k] =
            scf && scf->type == Json::ARR && scf->arr.size() == 3 ? (float)scf->arr[k].num_or(1.0) : 1.0f;
    m.specular_color_and_weight[3] = spec ? (float)spec->num_at("specularFactor", 1.0) : 1.0f;
    m.specular_roughness = spec_rough;
    m.specular_roughness_anisotropy = aniso;
    m.specular_haze = haze;
    m.specular_haze_spread = haze_spread;
    const Json* ior = ext ? ext->get("KHR_materials_ior") : nullptr;
    m.specular_ior = ior ? (float)ior->num_at("ior", 1.5) : 1.5f;
    m.specular_retroreflectivity = retro;
    m.normal_scale = ntex ? (float)ntex->num_at("scale", 1.0) : 1.0f;
    const Json* am = mat ? mat->get("alphaMode") : nullptr;
    const std::string mode = am && am->type == Json::STR ? am->str : "OPAQUE";
    m.alpha_mode = mode == "MASK" ? PT_ALPHA_MODE_MASK : mode == "BLEND" ? PT_ALPHA_MODE_BLEND : PT_ALPHA_MODE_OPAQUE;
    m.alpha_cutoff = mat ? (float)mat->num_at("alphaCutoff", 0.5) : 0.5f;
    m.pad = 0.0f;
    return m;
}

bool to_rgba8(const DecodedImage& img, std::vector<uint8_t>& rgba) {
    const size_t n = (size_t)img.width * img.height;
    rgba.assign(n * 4, 255);
    auto s8 = [&](size_t i, int c) -> uint8_t {
        uint16_t v = img.samples[i * img.channels + c];
        return img.bits == 16 ? (uint8_t)(v >> 8) : (uint8_t)v;
    };
    for (size_t i = 0; i < n; ++i) {
        uint8_t* o = &rgba[i * 4];
        switch (img.channels) {
            case 1: o[0] = o[1] = o[2] = s8(i, 0); break;
            case 2: o[0] = s8(i, 0); o[1] = s8(i, 1); o[2] = 0; break;
            case 3: for (int c = 0; c < 3; ++c) o[c] = s8(i, c); break;
            default: for (int c = 0; c < 4; ++c) o[c] = s8(i, c); break;
        }
    }
    return true;
}

}  // namespace

bool load_scene(const uint8_t* data, size_t size, Assets& A, std::string& err) {
    Doc d;
    std::vector<uint8_t> bin;
    bool have_bin = false;
    if (size >= 12 && std::memcmp(data, "glTF", 4) == 0) {
        uint32_t version, length;
        std::memcpy(&version, data + 4, 4);
        std::memcpy(&length, data + 8, 4);
        if (version != 2) return fail(err, "unsupported glTF binary version");
        if (length > size) return fail(err, "truncated .glb");
        size_t pos = 12;
        bool have_json = false;
        while (pos + 8 <= length) {
            uint32_t clen, ctype;
            std::memcpy(&clen, data + pos, 4);
            std::memcpy(&ctype, data + pos + 4, 4);
            if (clen > length - pos - 8) return fail(err, "truncated .glb chunk");
            const uint8_t* c = data + pos + 8;
            if (ctype == 0x4E4F534Au && !have_json) {
                JsonParser jp((const char*)c, clen);
                if (!jp.parse(d.j, err)) return false;
                have_json = true;
            } else if (ctype == 0x004E4942u && !have_bin) {
                bin.assign(c, c + clen);
                have_bin = true;
            }
            pos += 8 + (size_t)clen;
        }
        if (!have_json) return fail(err, ".glb without a JSON chunk");
    } else {
        JsonParser jp((const char*)data, size);
        if (!jp.parse(d.j, err)) return false;
    }
    if (d.j.type != Json::OBJ) return fail(err, "glTF root is not an object");
    if (const Json* buffers = d.j.get("buffers")) {
        for (size_t i = 0; i < buffers->arr.size(); ++i) {
            const Json* uri = buffers->arr[i].get("uri");
            std::vector<uint8_t> b;
            if (uri && uri->type == Json::STR) {
                if (!data_uri(uri->str, b, err)) return false;
            } else if (i == 0 && have_bin) {
                b = bin;
            } else {
                return fail(err, "buffer without data");
            }
            if ((double)b.size() < buffers->arr[i].num_at("byteLength", 0))
                return fail(err, "buffer shorter than its byteLength");
            d.buffers.push_back(std::move(b));
        }
    }
    // images -> atlases (world.rs:898-1075)
    std::set<long> srgb_images;
    auto note_srgb = [&](const Json* texinfo) {
        const Json* texs = d.j.get("textures");
        if (!texinfo || !texs) return;
        long ti = texinfo->index_at("index");
        if (ti >= 0 && (size_t)ti < texs->arr.size()) srgb_images.insert(texs->arr[ti].index_at("source"));
    };
    if (const Json* mats = d.j.get("materials"))
        for (const Json& m : mats->arr) {
            const Json* pbr = m.get("pbrMetallicRoughness");
            note_srgb(pbr ? pbr->get("baseColorTexture") : nullptr);
            note_srgb(m.get("emissiveTexture"));
            const Json* ext = m.get("extensions");
            const Json* spec = ext ? ext->get("KHR_materials_specular") : nullptr;
            note_srgb(spec ? spec->get("specularColorTexture") : nullptr);
        }
    std::vector<AtlasLayer> srgb_layers, lin_layers;
    std::map<long, ImageSlot> srgb_slots, lin_slots;
    if (const Json* images = d.j.get("images")) {
        for (size_t ii = 0; ii < images->arr.size(); ++ii) {
            const Json& im = images->arr[ii];
            std::vector<uint8_t> owned;
            const uint8_t* p = nullptr;
            size_t n = 0;
            if (const Json* uri = im.get("uri")) {
                if (!data_uri(uri->str, owned, err)) return false;
                p = owned.data();
                n = owned.size();
            } else {
                View v;
                if (!buffer_view(d, im.index_at("bufferView"), v, err)) return false;
                p = v.p;
                n = v.len;
            }
            DecodedImage img;
            const bool jpeg = n >= 3 && p[0] == 0xFF && p[1] == 0xD8 && p[2] == 0xFF;
            if (n >= 12 && std::memcmp(p, "RIFF", 4) == 0 && std::memcmp(p + 8, "WEBP", 4) == 0)
                return fail(err, "image " + std::to_string(ii) + " is WebP, which this loader does not decode (use PNG or JPEG)");
            if (!(jpeg ? decode_jpeg(p, n, img, err) : decode_png(p, n, img, err)))
                return fail(err, "image " + std::to_string(ii) + ": " + err);
            if (img.width > ATLAS || img.height > ATLAS) return fail(err, "single image is larger than the 8192 atlas");
            std::vector<uint8_t> rgba;
            to_rgba8(img, rgba);
            const bool is_srgb = srgb_images.count((long)ii) != 0;
            std::vector<AtlasLayer>& layers = is_srgb ? srgb_layers : lin_layers;
            std::vector<uint8_t>& pixels = is_srgb ? A.srgb : A.linear;
            uint32_t x = 0, y = 0;
            size_t li = 0;
            for (; li < layers.size(); ++li)
                if (layers[li].place(img.width, img.height, x, y)) break;
            if (li == layers.size()) {
                layers.emplace_back();
                pixels.resize(pixels.size() + (size_t)ATLAS * ATLAS * 4, 0);
                layers.back().place(img.width, img.height, x, y);
            }
            uint8_t* layer_px = pixels.data() + li * (size_t)ATLAS * ATLAS * 4;
            for (uint32_t row = 0; row < img.height; ++row)
                std::memcpy(layer_px + (((size_t)(y + row) * ATLAS) + x) * 4, &rgba[(size_t)row * img.width * 4],
                            (size_t)img.width * 4);
            ImageSlot slot;
            slot.layer = (int)li;
            slot.uv[0] = (float)x / (float)ATLAS;
            slot.uv[1] = (float)y / (float)ATLAS;
            slot.uv[2] = (float)img.width / (float)ATLAS;
            slot.uv[3] = (float)img.height / (float)ATLAS;
            (is_srgb ? srgb_slots : lin_slots)[(long)ii] = slot;
        }
    }
    A.srgb_layers = (uint32_t)srgb_layers.size();
    A.linear_layers = (uint32_t)lin_layers.size();

    // flatten the node hierarchy (world.rs:1076-1337)
    const Json* nodes = d.j.get("nodes");
    const Json* meshes = d.j.get("meshes");
    const Json* materials = d.j.get("materials");
    const size_t node_count = nodes && nodes->type == Json::ARR ? nodes->arr.size() : 0;
    std::vector<long> roots;
    const Json* scenes = d.j.get("scenes");
    const long scene_index = d.j.index_at("scene");
    if (scene_index >= 0 && scenes && (size_t)scene_index < scenes->arr.size()) {
        if (const Json* sn = scenes->arr[scene_index].get("nodes"))
            for (const Json& n : sn->arr) roots.push_back((long)n.num_or(-1));
    } else {
        for (size_t i = 0; i < node_count; ++i) roots.push_back((long)i);
    }
    std::vector<std::pair<long, M4>> stack;
    for (long r : roots) stack.emplace_back(r, identity());
    std::map<long, uint32_t> material_cache;
    size_t visits = 0;
    static const float DUMMY_UV[3][2] = {{0.0f, 0.0f}, {1.0f, 0.0f}, {0.0f, 1.0f}};
    while (!stack.empty()) {
        auto [ni, parent] = stack.back();
        stack.pop_back();
        if (ni < 0 || (size_t)ni >= node_count) return fail(err, "node index out of range");
        if (++visits > node_count * 64 + 1024) return fail(err, "node hierarchy has a cycle");
        const Json& node = nodes->arr[ni];
        M4 local = identity();
        if (const Json* mat = node.get("matrix")) {
            if (mat->type != Json::ARR || mat->arr.size() != 16) return fail(err, "bad node matrix");
            for (int k = 0; k < 16; ++k) local.c[k / 4][k % 4] = (float)mat->arr[k].num_or(0.0);
        } else {
            float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
            if (const Json* a = node.get("translation"))
                for (size_t k = 0; k < 3 && k < a->arr.size(); ++k) t[k] = (float)a->arr[k].num_or(0.0);
            if (const Json* a = node.get("rotation"))
                for (size_t k = 0; k < 4 && k < a->arr.size(); ++k) q[k] = (float)a->arr[k].num_or(0.0);
            if (const Json* a = node.get("scale"))
                for (size_t k = 0; k < 3 && k < a->arr.size(); ++k) s[k] = (float)a->arr[k].num_or(1.0);
            M4 T = identity(), S = identity();
            for (int k = 0; k < 3; ++k) {
                T.c[3][k] = t[k];
                S.c[k][k] = s[k];
            }
            local = mul(mul(T, from_quat(q[0], q[1], q[2], q[3])), S);
        }
        const M4 model = mul(parent, local);
        const long mi = node.index_at("mesh");
        if (mi >= 0) {
            if (!meshes || (size_t)mi >= meshes->arr.size()) return fail(err, "mesh index out of range");
            const M4 normal_mat = transpose(inverse(model));
            const Json* prims = meshes->arr[mi].get("primitives");
            for (size_t pi = 0; prims && pi < prims->arr.size(); ++pi) {
                const Json& prim = prims->arr[pi];
                if ((long)prim.num_at("mode", 4) != 4) {
                    A.warnings.push_back("mesh mode is not triangles, skipping primitive");
                    continue;
                }
                const Json* attrs = prim.get("attributes");
                const long pos_a = attrs ? attrs->index_at("POSITION") : -1;
                if (pos_a < 0) {
                    A.warnings.push_back("skipping primitive with no vertex positions");
                    continue;
                }
                Accessor ap;
                if (!accessor(d, pos_a, ap, err, 3)) return false;
                const size_t nv = ap.count;
                std::vector<float> pos(nv * 3);
                for (size_t v = 0; v < nv; ++v) {
                    float p[3] = {ap.f32(v, 0), ap.f32(v, 1), ap.f32(v, 2)};
                    point3(model, p, &pos[3 * v]);
                }
                std::vector<uint32_t> idx;
                if (const long ia = prim.index_at("indices"); ia >= 0) {
                    Accessor ai;
                    if (!accessor(d, ia, ai, err, 1)) return false;
                    idx.resize(ai.count);
                    for (size_t k = 0; k < ai.count; ++k) {
                        idx[k] = ai.u32(k);
                        if (idx[k] >= nv) return fail(err, "vertex index out of range");
                    }
                } else {
                    idx.resize(nv);
                    for (size_t k = 0; k < nv; ++k) idx[k] = (uint32_t)k;
                }
                idx.resize(idx.size() / 3 * 3);
                std::vector<float> uv(nv * 2);
                if (const long ta = attrs->index_at("TEXCOORD_0"); ta >= 0) {
                    Accessor at;
                    if (!accessor(d, ta, at, err, 2) || at.count != nv) return fail(err, err.empty() ? "TEXCOORD_0 count" : err);
                    for (size_t v = 0; v < nv; ++v) {
                        uv[2 * v] = at.f32(v, 0);
                        uv[2 * v + 1] = at.f32(v, 1);
                    }
                } else {
                    for (size_t v = 0; v < nv; ++v) {
                        uv[2 * v] = DUMMY_UV[v % 3][0];
                        uv[2 * v + 1] = DUMMY_UV[v % 3][1];
                    }
                }
                std::vector<float> nrm(nv * 3, 0.0f), tan_v;
                bool computed = false;
                if (const long na = attrs->index_at("NORMAL"); na >= 0) {
                    Accessor an;
                    if (!accessor(d, na, an, err, 3) || an.count != nv) return fail(err, err.empty() ? "NORMAL count" : err);
                    for (size_t v = 0; v < nv; ++v) {
                        float n[3] = {an.f32(v, 0), an.f32(v, 1), an.f32(v, 2)};
                        vector3(normal_mat, n, &nrm[3 * v]);
                        normalize3(&nrm[3 * v]);
                    }
                    if (const long tga = attrs->index_at("TANGENT"); tga >= 0) {
                        Accessor at;
                        if (!accessor(d, tga, at, err, 4) || at.count != nv) return fail(err, err.empty() ? "TANGENT count" : err);
                        for (size_t v = 0; v < nv && !computed; ++v) {
                            float t[4] = {at.f32(v, 0), at.f32(v, 1), at.f32(v, 2), at.f32(v, 3)};
                            if (t[0] == 1.0f && t[1] == 0.0f && t[2] == 0.0f && t[3] == 1.0f) {
                                computed = true;  // a placeholder tangent
                                break;
                            }
                            float o[3];
                            vector3(model, t, o);
                            normalize3(o);
                            tan_v.insert(tan_v.end(), {o[0], o[1], o[2], t[3]});
                        }
                        for (size_t k = 0; !computed && k + 2 < idx.size(); k += 3) {
                            const float w0 = tan_v[4 * idx[k] + 3], w1 = tan_v[4 * idx[k + 1] + 3], w2 = tan_v[4 * idx[k + 2] + 3];
                            if (w0 != w1 || w0 != w2) computed = true;
                        }
                    } else {
                        computed = true;
                    }
                } else {
                    for (size_t k = 0; k + 2 < idx.size(); k += 3) {
                        const float* p0 = &pos[3 * idx[k]];
                        const float* p1 = &pos[3 * idx[k + 1]];
                        const float* p2 = &pos[3 * idx[k + 2]];
                        const float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
                        const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
                        float n[3] = {a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]};
                        normalize3(n);
                        for (int c = 0; c < 3; ++c) std::memcpy(&nrm[3 * idx[k + c]], n, sizeof(n));
                    }
                    computed = true;
                }
                std::vector<float> corner_tan;
                if (computed) {
                    MeshView mv{idx.data(), idx.size(), pos.data(), nrm.data(), uv.data()};
                    generate_tangents(mv, corner_tan);
                    for (size_t k = 3; k < corner_tan.size(); k += 4) corner_tan[k] = -corner_tan[k];
                }
                // material (created at first use, shared per glTF material)
                uint32_t mat_index;
                const long gm = prim.index_at("material");
                if (gm >= 0) {
                    if (!materials || (size_t)gm >= materials->arr.size()) return fail(err, "material index out of range");
                    auto it = material_cache.find(gm);
                    if (it == material_cache.end()) {
                        mat_index = (uint32_t)A.materials.size();
                        A.materials.push_back(make_material(&materials->arr[gm], srgb_slots, lin_slots, d, A.warnings));
                        material_cache[gm] = mat_index;
                    } else {
                        mat_index = it->second;
                    }
                } else {
                    mat_index = (uint32_t)A.materials.size();
                    A.materials.push_back(make_material(nullptr, srgb_slots, lin_slots, d, A.warnings));
                }
                for (size_t k = 0, tri = 0; k + 2 < idx.size(); k += 3, ++tri) {
                    PtTriangleGeometry g;
                    PtTriangleAttribute at{};
                    at.material = mat_index;
                    for (int c = 0; c < 3; ++c) {
                        const uint32_t v = idx[k + c];
                        std::memcpy(&g.v[3 * c], &pos[3 * v], 12);
                        std::memcpy(&at.n[3 * c], &nrm[3 * v], 12);
                        std::memcpy(&at.uv[2 * c], &uv[2 * v], 8);
                        float* t = c == 0 ? at.t0 : c == 1 ? at.t1 : at.t2;
                        if (computed) std::memcpy(t, &corner_tan[4 * (3 * tri + c)], 16);
                        else std::memcpy(t, &tan_v[4 * v], 16);
                    }
                    A.geometries.push_back(g);
                    A.attributes.push_back(at);
                }
            }
        }
        if (const Json* ch = node.get("children"))
            for (const Json& c : ch->arr) stack.emplace_back((long)c.num_or(-1), model);
    }
    if (A.geometries.empty()) return fail(err, "scene has no triangles");
    return true;
}

}  // namespace ptcore

using namespace ptcore;

struct pt_assets : ptcore::Assets {};

extern "C" {

pt_status pt_assets_create(pt_assets** out) {
    if (!out) return set_error(PT_ERR_INVALID, "pt_assets_create: null argument");
    *out = new (std::nothrow) pt_assets();
    return *out ? PT_OK : set_error(PT_ERR_MEMORY, "out of host memory");
}

pt_status pt_assets_destroy(pt_assets* a) {
    delete a;
    return PT_OK;
}

pt_status pt_assets_load_scene(pt_assets* a, const uint8_t* bytes, size_t size) {
    if (!a || (!bytes && size)) return set_error(PT_ERR_INVALID, "pt_assets_load_scene: null argument");
    Assets fresh;
    std::string err;
    try {
        if (!load_scene(bytes, size, fresh, err)) return set_error(PT_ERR_INVALID, "scene: %s", err.c_str());
    } catch (const std::bad_alloc&) {
        return set_error(PT_ERR_MEMORY, "out of host memory while loading the scene");
    }
    for (std::string& w : fresh.warnings) w = utf8_clean(w.c_str());  // they can quote the file's JSON
    fresh.sky = std::move(a->sky);
    static_cast<Assets&>(*a) = std::move(fresh);
    return PT_OK;
}

pt_status pt_assets_load_skybox(pt_assets* a, const uint8_t* bytes, size_t size) {
    if (!a || (!bytes && size)) return set_error(PT_ERR_INVALID, "pt_assets_load_skybox: null argument");
    FloatImage sky;
    std::string err;
    try {
        if (!decode_sky_image(bytes, size, sky, err)) return set_error(PT_ERR_INVALID, "skybox: %s", err.c_str());
    } catch (const std::bad_alloc&) {
        return set_error(PT_ERR_MEMORY, "out of host memory while decoding the skybox");
    }
    a->sky = std::move(sky);
    return PT_OK;
}

static pt_status read_file(const char* path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(PT_ERR_INVALID, "cannot open '%s'", path);
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return set_error(PT_ERR_INVALID, "cannot read '%s'", path);
    }
    out.resize((size_t)n);
    size_t got = n ? std::fread(out.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    if (got != (size_t)n) return set_error(PT_ERR_INVALID, "short read of '%s'", path);
    return PT_OK;
}

pt_status pt_assets_load_scene_file(pt_assets* a, const char* path) {
    if (!a || !path) return set_error(PT_ERR_INVALID, "pt_assets_load_scene_file: null argument");
    std::vector<uint8_t> bytes;
    pt_status st = read_file(path, bytes);
    return st ? st : pt_assets_load_scene(a, bytes.data(), bytes.size());
}

pt_status pt_assets_load_skybox_file(pt_assets* a, const char* path) {
    if (!a || !path) return set_error(PT_ERR_INVALID, "pt_assets_load_skybox_file: null argument");
    std::vector<uint8_t> bytes;
    pt_status st = read_file(path, bytes);
    return st ? st : pt_assets_load_skybox(a, bytes.data(), bytes.size());
}

pt_status pt_assets_info(const pt_assets* a, pt_assets_summary* out) {
    if (!a || !out) return set_error(PT_ERR_INVALID, "pt_assets_info: null argument");
    out->triangle_count = (uint32_t)a->geometries.size();
    out->material_count = (uint32_t)a->materials.size();
    out->srgb_atlas_count = a->srgb_layers;
    out->linear_atlas_count = a->linear_layers;
    out->atlas_size = (a->srgb_layers || a->linear_layers) ? ATLAS : 0u;
    out->skybox_width = a->sky.width;
    out->skybox_height = a->sky.height;
    out->warning_count = (uint32_t)a->warnings.size();
    return PT_OK;
}

const char* pt_assets_warning(const pt_assets* a, uint32_t i) {
    return a && i < a->warnings.size() ? a->warnings[i].c_str() : "";
}

pt_status pt_assets_world_desc(const pt_assets* a, pt_world_desc* d) {
    if (!a || !d) return set_error(PT_ERR_INVALID, "pt_assets_world_desc: null argument");
    if (a->geometries.empty()) return set_error(PT_ERR_STATE, "no scene loaded");
    if (a->sky.rgba.empty()) return set_error(PT_ERR_STATE, "no skybox loaded");
    std::memset(d, 0, sizeof(*d));
    d->geometries = a->geometries.data();
    d->attributes = a->attributes.data();
    d->triangle_count = (uint32_t)a->geometries.size();
    d->materials = a->materials.data();
    d->material_count = (uint32_t)a->materials.size();
    d->srgb_atlases = a->srgb.empty() ? nullptr : a->srgb.data();
    d->srgb_atlas_count = a->srgb_layers;
    d->linear_atlases = a->linear.empty() ? nullptr : a->linear.data();
    d->linear_atlas_count = a->linear_layers;
    d->atlas_size = (a->srgb_layers || a->linear_layers) ? ATLAS : 0u;
    d->skybox_rgba = a->sky.rgba.data();
    d->skybox_width = a->sky.width;
    d->skybox_height = a->sky.height;
    d->skybox_luminance = -1.0f;
    const float nan = std::numeric_limits<float>::quiet_NaN();
    for (int k = 0; k < 3; ++k) d->camera_position[k] = nan;
    d->camera_yaw_degrees = d->camera_pitch_degrees = d->camera_fov_degrees = nan;
    d->aspect_ratio = 1.0f;
    return PT_OK;
}

}  // extern "C"
