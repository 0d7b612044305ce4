// scene_export.cpp - writing worlds in formats the reference reads, and the
// built device arrays as a .ptscene file.
//
// pt_world_export_glb writes a world's input (the triangle soup, materials,
// atlases and sky that pt_world_build takes) as a glTF binary and a sky
// image, so that the reference's Scene::new and Skybox::new (world.rs:887-1423,
// 1872-1873) and this library's loader rebuild the same input:
//   * one node, no transform, one mesh; one primitive per run of consecutive
//     triangles with one material, so the loader's flattening keeps the
//     triangle order; unindexed POSITION, NORMAL, TANGENT and TEXCOORD_0 per
//     corner;
//   * glTF material i is world material i (the loader numbers materials in
//     order of first use, so the world's must be numbered that way);
//   * every texture rectangle of the atlases becomes an RGBA PNG image; sRGB
//     ones, then linear ones, each in atlas order (layer, y, x), which must be
//     the order in which the loader's packer places them back;
//   * the OpenPBR parameters go in the material extras, the rest in the core
//     glTF fields and KHR_materials_{specular, ior, emissive_strength};
//   * the sky as an uncompressed FLOAT OpenEXR, which round-trips RGBA32F, or,
//     for a path ending in .hdr, as Radiance RGBE, refused unless every value
//     decodes back exactly.
// The camera and the sky luminance are not part of either file: they are the
// reference's render flags (headless.rs:14-64).
//
// No counterpart in the reference for .ptscene: a flat dump of a
// pt_scene_desc (the device-layout arrays, BVH included) for tools that
// want the identical serialized BVH.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pt_assets.h"
#include "gltf_common.h"
#include "image_io.h"
#include "pt_internal.h"

namespace ptcore {
namespace {

void jnum(std::string& o, double v) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.9g", v);
    o += b;
}
void jarr(std::string& o, const float* v, int n) {
    o += '[';
    for (int i = 0; i < n; ++i) {
        if (i) o += ',';
        jnum(o, v[i]);
    }
    o += ']';
}

struct Bin {
    std::vector<uint8_t> data;
    std::string views;  // JSON bufferViews
    uint32_t count = 0;
    uint32_t add(const void* p, size_t n, const char* target) {
        while (data.size() % 4) data.push_back(0);
        const size_t at = data.size();
        data.insert(data.end(), (const uint8_t*)p, (const uint8_t*)p + n);
        if (count) views += ',';
        views += "{\"buffer\":0,\"byteOffset\":" + std::to_string(at) + ",\"byteLength\":" + std::to_string(n);
        if (target) views += std::string(",\"target\":") + target;
        views += '}';
        return count++;
    }
};

// a texture rectangle of an atlas layer, in texels
struct TexRect {
    bool srgb;
    int layer;
    uint32_t x, y, w, h;
    bool operator<(const TexRect& o) const {
        return std::tie(srgb, layer, y, x, w, h) < std::tie(o.srgb, o.layer, o.y, o.x, o.w, o.h);
    }
};

bool rect_of(const pt_world_desc* d, bool srgb, int layer, const float uv[4], TexRect& r, std::string& err) {
    const double s = d->atlas_size;
    const double px[4] = {uv[0] * s, uv[1] * s, uv[2] * s, uv[3] * s};
    for (double v : px)
        if (v < 0 || v != std::floor(v)) {
            err = "a texture's atlas rectangle is not whole texels";
            return false;
        }
    r = {srgb, layer, (uint32_t)px[0], (uint32_t)px[1], (uint32_t)px[2], (uint32_t)px[3]};
    // the loader writes uv = texel / 8192 in float
    const float back[4] = {(float)r.x / (float)ATLAS, (float)r.y / (float)ATLAS, (float)r.w / (float)ATLAS,
                           (float)r.h / (float)ATLAS};
    if (std::memcmp(back, uv, sizeof(back)) != 0 || r.w == 0 || r.h == 0 || r.x + r.w > ATLAS || r.y + r.h > ATLAS) {
        err = "a texture's atlas offset or scale is not what the loader would give it";
        return false;
    }
    return true;
}

bool export_glb(const pt_world_desc* d, std::vector<uint8_t>& glb, uint32_t& inexact, std::string& err) {
    inexact = 0;
    if (!d->triangle_count || !d->geometries || !d->attributes || !d->material_count || !d->materials) {
        err = "world has no triangles or materials";
        return false;
    }
    if ((d->srgb_atlas_count || d->linear_atlas_count) && d->atlas_size != ATLAS) {
        err = "atlas layers must be 8192^2, as the loader packs them";
        return false;
    }
    // texture rectangles, in the order the loader must place them back
    std::map<TexRect, int> images;
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const PtMaterial& m = d->materials[i];
        struct Slot {
            int32_t layer;
            const float* uv;
            bool srgb;
        } slots[6] = {{m.base_color_tex_layer, m.base_color_uv, true},
                      {m.emission_tex_layer, m.emission_uv, true},
                      {m.specular_color_tex_layer, m.specular_color_uv, true},
                      {m.metallic_roughness_tex_layer, m.metallic_roughness_uv, false},
                      {m.normal_tex_layer, m.normal_uv, false},
                      {m.specular_weight_tex_layer, m.specular_weight_uv, false}};
        for (const Slot& sl : slots) {
            if (sl.layer < 0) continue;
            if ((uint32_t)sl.layer >= (sl.srgb ? d->srgb_atlas_count : d->linear_atlas_count)) {
                err = "a material names a missing atlas layer";
                return false;
            }
            TexRect r;
            if (!rect_of(d, sl.srgb, sl.layer, sl.uv, r, err)) return false;
            images[r] = 0;
        }
    }
    {  // the loader's packer must put every image back where it is
        int k = 0;
        std::vector<AtlasLayer> packers[2];
        for (auto& kv : images) {
            kv.second = k++;
            const TexRect& r = kv.first;
            std::vector<AtlasLayer>& L = packers[r.srgb ? 0 : 1];
            uint32_t x = 0, y = 0;
            size_t li = 0;
            for (; li < L.size(); ++li)
                if (L[li].place(r.w, r.h, x, y)) break;
            if (li == L.size()) {
                L.emplace_back();
                L.back().place(r.w, r.h, x, y);
            }
            if ((int)li != r.layer || x != r.x || y != r.y) {
                err = "the atlas layout is not the one the loader's packer gives the textures";
                return false;
            }
        }
    }
    Bin bin;
    std::string accessors, meshes_prims, json_images, json_textures;
    uint32_t n_accessors = 0;
    auto accessor = [&](uint32_t view, uint32_t count, const char* type, const float* mn, const float* mx, int ncomp) {
        if (n_accessors) accessors += ',';
        accessors += "{\"bufferView\":" + std::to_string(view) + ",\"componentType\":5126,\"count\":" +
                     std::to_string(count) + ",\"type\":\"" + type + "\"";
        if (mn) {
            accessors += ",\"min\":";
            jarr(accessors, mn, ncomp);
            accessors += ",\"max\":";
            jarr(accessors, mx, ncomp);
        }
        accessors += '}';
        return n_accessors++;
    };
    // images (PNG in the binary chunk)
    for (const auto& kv : images) {
        const TexRect& r = kv.first;
        const uint8_t* layer = (r.srgb ? d->srgb_atlases : d->linear_atlases) + (size_t)r.layer * ATLAS * ATLAS * 4;
        std::vector<uint8_t> png;
        if (!encode_png_channels(layer + ((size_t)r.y * ATLAS + r.x) * 4, (size_t)ATLAS * 4, r.w, r.h, 4, png, err))
            return false;
        const uint32_t v = bin.add(png.data(), png.size(), nullptr);
        if (kv.second) json_images += ',', json_textures += ',';
        json_images += "{\"bufferView\":" + std::to_string(v) + ",\"mimeType\":\"image/png\"}";
        json_textures += "{\"source\":" + std::to_string(kv.second) + "}";
    }
    // primitives: runs of triangles with one material
    uint32_t n_prims = 0;
    for (uint32_t t0 = 0; t0 < d->triangle_count;) {
        const uint32_t mat = d->attributes[t0].material;
        if (mat >= d->material_count) {
            err = "a triangle names a missing material";
            return false;
        }
        uint32_t t1 = t0 + 1;
        while (t1 < d->triangle_count && d->attributes[t1].material == mat) ++t1;
        const uint32_t nv = 3 * (t1 - t0);
        std::vector<float> pos((size_t)nv * 3), nrm((size_t)nv * 3), tan((size_t)nv * 4), uv((size_t)nv * 2);
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t t = t0; t < t1; ++t) {
            const PtTriangleAttribute& at = d->attributes[t];
            const float* tg[3] = {at.t0, at.t1, at.t2};
            for (int c = 0; c < 3; ++c) {
                const size_t v = (size_t)3 * (t - t0) + c;
                for (int k = 0; k < 3; ++k) {
                    pos[3 * v + k] = d->geometries[t].v[3 * c + k];
                    mn[k] = std::min(mn[k], pos[3 * v + k]);
                    mx[k] = std::max(mx[k], pos[3 * v + k]);
                }
                // directions that load back as exactly the stored ones
                if (!unit_preimage(&at.n[3 * c], &nrm[3 * v])) {
                    std::memcpy(&nrm[3 * v], &at.n[3 * c], 12);
                    ++inexact;
                }
                if (!unit_preimage(tg[c], &tan[4 * v])) {
                    std::memcpy(&tan[4 * v], tg[c], 12);
                    ++inexact;
                }
                tan[4 * v + 3] = tg[c][3];
                // the loader takes a stored (1,0,0,1) for a placeholder; (2,0,0,1) loads as (1,0,0,1)
                if (tan[4 * v] == 1.0f && tan[4 * v + 1] == 0.0f && tan[4 * v + 2] == 0.0f && tan[4 * v + 3] == 1.0f)
                    tan[4 * v] = 2.0f;
                std::memcpy(&uv[2 * v], &at.uv[2 * c], 8);
            }
        }
        for (float f : pos)
            if (!std::isfinite(f)) {
                err = "a vertex position is not finite";
                return false;
            }
        const uint32_t ap = accessor(bin.add(pos.data(), pos.size() * 4, "34962"), nv, "VEC3", mn, mx, 3);
        const uint32_t an = accessor(bin.add(nrm.data(), nrm.size() * 4, "34962"), nv, "VEC3", nullptr, nullptr, 0);
        const uint32_t atg = accessor(bin.add(tan.data(), tan.size() * 4, "34962"), nv, "VEC4", nullptr, nullptr, 0);
        const uint32_t auv = accessor(bin.add(uv.data(), uv.size() * 4, "34962"), nv, "VEC2", nullptr, nullptr, 0);
        if (n_prims++) meshes_prims += ',';
        meshes_prims += "{\"attributes\":{\"POSITION\":" + std::to_string(ap) + ",\"NORMAL\":" + std::to_string(an) +
                        ",\"TANGENT\":" + std::to_string(atg) + ",\"TEXCOORD_0\":" + std::to_string(auv) +
                        "},\"material\":" + std::to_string(mat) + ",\"mode\":4}";
        t0 = t1;
    }
    // materials
    std::string mats;
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const PtMaterial& m = d->materials[i];
        auto texref = [&](bool srgb, int32_t layer, const float* uv, std::string& out) {
            TexRect r;
            if (!rect_of(d, srgb, layer, uv, r, err)) return false;
            out = "{\"index\":" + std::to_string(images.at(r)) + "}";
            return true;
        };
        {  // the glTF schema's ranges for the core and KHR factors
            auto in01 = [](float v) { return v >= 0.0f && v <= 1.0f; };
            const char* bad = nullptr;
            for (int k = 0; k < 4; ++k)
                if (!in01(m.base_color[k])) bad = "baseColorFactor";
            for (int k = 0; k < 3; ++k) {
                if (!in01(m.emission_color_and_luminance[k])) bad = "emissiveFactor";
                if (!(m.specular_color_and_weight[k] >= 0.0f) || std::isinf(m.specular_color_and_weight[k]))
                    bad = "specularColorFactor";
            }
            if (!in01(m.base_metalness)) bad = "metallicFactor";
            if (!in01(m.base_diffuse_roughness)) bad = "roughnessFactor (the diffuse roughness)";
            if (!in01(m.specular_color_and_weight[3])) bad = "specularFactor";
            if (!(m.emission_color_and_luminance[3] >= 0.0f) || std::isinf(m.emission_color_and_luminance[3]))
                bad = "emissiveStrength";
            if (!(m.specular_ior == 0.0f || (m.specular_ior >= 1.0f && !std::isinf(m.specular_ior)))) bad = "ior";
            if (!(m.alpha_cutoff >= 0.0f) || std::isinf(m.alpha_cutoff)) bad = "alphaCutoff";
            if (bad) {
                err = "material " + std::to_string(i) + ": " + bad + " outside its glTF range cannot be written";
                return false;
            }
        }
        if (i) mats += ',';
        mats += "{\"pbrMetallicRoughness\":{\"baseColorFactor\":";
        jarr(mats, m.base_color, 4);
        mats += ",\"metallicFactor\":";
        jnum(mats, m.base_metalness);
        mats += ",\"roughnessFactor\":";
        jnum(mats, m.base_diffuse_roughness);
        std::string ref;
        if (m.base_color_tex_layer >= 0) {
            if (!texref(true, m.base_color_tex_layer, m.base_color_uv, ref)) return false;
            mats += ",\"baseColorTexture\":" + ref;
        }
        if (m.metallic_roughness_tex_layer >= 0) {
            if (!texref(false, m.metallic_roughness_tex_layer, m.metallic_roughness_uv, ref)) return false;
            mats += ",\"metallicRoughnessTexture\":" + ref;
        }
        mats += "},\"emissiveFactor\":";
        jarr(mats, m.emission_color_and_luminance, 3);
        if (m.emission_tex_layer >= 0) {
            if (!texref(true, m.emission_tex_layer, m.emission_uv, ref)) return false;
            mats += ",\"emissiveTexture\":" + ref;
        }
        if (m.normal_tex_layer >= 0) {
            if (!texref(false, m.normal_tex_layer, m.normal_uv, ref)) return false;
            ref.pop_back();
            mats += ",\"normalTexture\":" + ref + ",\"scale\":";
            jnum(mats, m.normal_scale);
            mats += '}';
        } else if (m.normal_scale != 1.0f) {
            err = "a normal scale other than 1 without a normal texture cannot be written to glTF";
            return false;
        }
        mats += ",\"alphaMode\":\"";
        mats += m.alpha_mode == PT_ALPHA_MODE_MASK ? "MASK" : m.alpha_mode == PT_ALPHA_MODE_BLEND ? "BLEND" : "OPAQUE";
        mats += "\",\"alphaCutoff\":";
        jnum(mats, m.alpha_cutoff);
        mats += ",\"extensions\":{\"KHR_materials_emissive_strength\":{\"emissiveStrength\":";
        jnum(mats, m.emission_color_and_luminance[3]);
        mats += "},\"KHR_materials_ior\":{\"ior\":";
        jnum(mats, m.specular_ior);
        mats += "},\"KHR_materials_specular\":{\"specularFactor\":";
        jnum(mats, m.specular_color_and_weight[3]);
        mats += ",\"specularColorFactor\":";
        jarr(mats, m.specular_color_and_weight, 3);
        if (m.specular_weight_tex_layer >= 0) {
            if (!texref(false, m.specular_weight_tex_layer, m.specular_weight_uv, ref)) return false;
            mats += ",\"specularTexture\":" + ref;
        }
        if (m.specular_color_tex_layer >= 0) {
            if (!texref(true, m.specular_color_tex_layer, m.specular_color_uv, ref)) return false;
            mats += ",\"specularColorTexture\":" + ref;
        }
        mats += "}},\"extras\":{";
        const std::pair<const char*, float> extras[7] = {
            {"OpenPBR_base_weight", m.base_weight},
            {"OpenPBR_emission_weight", m.emission_weight},
            {"OpenPBR_specular_roughness", m.specular_roughness},
            {"OpenPBR_specular_roughness_anisotropy", m.specular_roughness_anisotropy},
            {"OpenPBR_specular_haze", m.specular_haze},
            {"OpenPBR_specular_haze_spread", m.specular_haze_spread},
            {"OpenPBR_specular_retroreflectivity", m.specular_retroreflectivity}};
        for (int k = 0; k < 7; ++k) {
            if (!(extras[k].second >= 0.0f && extras[k].second <= 1.0f)) {
                err = std::string(extras[k].first) + " outside [0, 1] cannot be written (the loader ignores it)";
                return false;
            }
            if (k) mats += ',';
            mats += std::string("\"") + extras[k].first + "\":";
            jnum(mats, extras[k].second);
        }
        mats += "}}";
    }
    std::string j = "{\"asset\":{\"version\":\"2.0\",\"generator\":\"ptcore pt_world_export_glb\"},"
                    "\"extensionsUsed\":[\"KHR_materials_emissive_strength\",\"KHR_materials_ior\","
                    "\"KHR_materials_specular\"],\"scene\":0,\"scenes\":[{\"nodes\":[0]}],\"nodes\":[{\"mesh\":0}],"
                    "\"meshes\":[{\"primitives\":[" + meshes_prims + "]}],\"materials\":[" + mats + "]";
    if (!images.empty()) j += ",\"images\":[" + json_images + "],\"textures\":[" + json_textures + "]";
    while (bin.data.size() % 4) bin.data.push_back(0);
    j += ",\"accessors\":[" + accessors + "],\"bufferViews\":[" + bin.views + "],\"buffers\":[{\"byteLength\":" +
         std::to_string(bin.data.size()) + "}]}";
    while (j.size() % 4) j += ' ';
    const uint64_t total = 12 + 8 + j.size() + 8 + bin.data.size();
    if (total > 0xffffffffull) {
        err = "scene too large for one .glb (4 GiB)";
        return false;
    }
    glb.clear();
    glb.reserve(total);
    auto u32 = [&](uint32_t v) {
        for (int k = 0; k < 4; ++k) glb.push_back((uint8_t)(v >> (8 * k)));
    };
    glb.insert(glb.end(), {'g', 'l', 'T', 'F'});
    u32(2);
    u32((uint32_t)total);
    u32((uint32_t)j.size());
    u32(0x4E4F534Au);
    glb.insert(glb.end(), j.begin(), j.end());
    u32((uint32_t)bin.data.size());
    u32(0x004E4942u);
    glb.insert(glb.end(), bin.data.begin(), bin.data.end());
    return true;
}

bool write_file(const char* path, const std::vector<uint8_t>& bytes, std::string& err) {
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("cannot write ") + path;
        return false;
    }
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    if (std::fclose(f) != 0 || !ok) {
        err = std::string("writing ") + path + " failed";
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ .ptscene
// "PTSCENE1", u32 section count, then sections of u32 tag, u32 zero, u64 size
// and the bytes, each padded to 16.
constexpr uint32_t tag(const char (&s)[5]) {
    return (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16) | ((uint32_t)s[3] << 24);
}

}  // namespace
}  // namespace ptcore

struct pt_scene_file {
    std::vector<PtTriangleGeometry> geo;
    std::vector<PtTriangleAttribute> attr;
    std::vector<PtMaterial> mat;
    std::vector<PtBvhNode> bvh;
    std::vector<uint8_t> srgb, linear;
    uint32_t srgb_count = 0, linear_count = 0, atlas_size = 0;
    PtCamera camera{};
    std::vector<float> sky;
    PtSkyboxUniforms skybox{};
    std::vector<PtSkyboxAliasEntry> sky_alias;
    PtSceneUniforms uniforms{};
    std::vector<PtSceneAliasEntry> scene_alias;
};

using namespace ptcore;

extern "C" {

pt_status pt_world_export_glb(const pt_world_desc* d, const char* glb_path, const char* sky_path,
                              uint32_t* inexact_directions) {
    if (!d || !glb_path) return set_error(PT_ERR_INVALID, "pt_world_export_glb: null argument");
    std::string err;
    std::vector<uint8_t> glb;
    uint32_t inexact = 0;
    try {
        if (!export_glb(d, glb, inexact, err) || !write_file(glb_path, glb, err))
            return set_error(PT_ERR_INVALID, "%s", err.c_str());
        if (inexact_directions) *inexact_directions = inexact;
        if (sky_path) {
            if (!d->skybox_rgba || !d->skybox_width || !d->skybox_height) return set_error(PT_ERR_INVALID, "world has no sky");
            const size_t n = std::strlen(sky_path);
            const bool hdr = n >= 4 && (std::strcmp(sky_path + n - 4, ".hdr") == 0 || std::strcmp(sky_path + n - 4, ".HDR") == 0);
            std::vector<uint8_t> img;
            const bool ok = hdr ? encode_hdr_exact(d->skybox_rgba, d->skybox_width, d->skybox_height, img, err)
                                : encode_exr_float(d->skybox_rgba, d->skybox_width, d->skybox_height, img, err);
            if (!ok || !write_file(sky_path, img, err)) return set_error(PT_ERR_INVALID, "%s", err.c_str());
        }
    } catch (const std::bad_alloc&) {
        return set_error(PT_ERR_MEMORY, "out of host memory while writing the scene");
    }
    return PT_OK;
}

pt_status pt_scene_save(const pt_scene_desc* d, const char* path) {
    if (!d || !path) return set_error(PT_ERR_INVALID, "pt_scene_save: null argument");
    const size_t layer = (size_t)d->atlas_size * d->atlas_size * 4;
    struct Sec {
        uint32_t tag;
        const void* p;
        uint64_t n;
    } secs[] = {
        {tag("GEOM"), d->geometries, (uint64_t)d->triangle_count * sizeof(PtTriangleGeometry)},
        {tag("ATTR"), d->attributes, (uint64_t)d->triangle_count * sizeof(PtTriangleAttribute)},
        {tag("MATL"), d->materials, (uint64_t)d->material_count * sizeof(PtMaterial)},
        {tag("BVHN"), d->bvh_nodes, (uint64_t)d->bvh_node_count * sizeof(PtBvhNode)},
        {tag("ASIZ"), &d->atlas_size, 4},
        {tag("SRGB"), d->srgb_atlases, (uint64_t)d->srgb_atlas_count * layer},
        {tag("LINR"), d->linear_atlases, (uint64_t)d->linear_atlas_count * layer},
        {tag("CAMR"), &d->camera, sizeof(PtCamera)},
        {tag("SKYU"), &d->skybox, sizeof(PtSkyboxUniforms)},
        {tag("SKYR"), d->skybox_rgba, (uint64_t)d->skybox.width * d->skybox.height * 16},
        {tag("SKYA"), d->skybox_alias, (uint64_t)d->skybox_alias_count * sizeof(PtSkyboxAliasEntry)},
        {tag("SCNU"), &d->scene_uniforms, sizeof(PtSceneUniforms)},
        {tag("SCNA"), d->scene_alias, (uint64_t)d->scene_alias_count * sizeof(PtSceneAliasEntry)},
    };
    for (const Sec& s : secs)
        if (s.n && !s.p) return set_error(PT_ERR_INVALID, "pt_scene_save: an array with entries has no data");
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(PT_ERR_INVALID, "cannot write %s", path);
    bool ok = std::fwrite("PTSCENE1", 1, 8, f) == 8;
    const uint32_t count = sizeof(secs) / sizeof(secs[0]);
    ok = ok && std::fwrite(&count, 4, 1, f) == 1;
    static const uint8_t zeros[16] = {0};
    ok = ok && std::fwrite(zeros, 1, 4, f) == 4;  // header padded to 16
    for (const Sec& s : secs) {
        const uint32_t hdr[2] = {s.tag, 0};
        ok = ok && std::fwrite(hdr, 4, 2, f) == 2 && std::fwrite(&s.n, 8, 1, f) == 1;
        ok = ok && (s.n == 0 || std::fwrite(s.p, 1, s.n, f) == s.n);
        ok = ok && std::fwrite(zeros, 1, (16 - s.n % 16) % 16, f) == (16 - s.n % 16) % 16;
    }
    if (std::fclose(f) != 0 || !ok) return set_error(PT_ERR_INVALID, "writing %s failed", path);
    return PT_OK;
}

pt_status pt_scene_file_load(const char* path, pt_scene_file** out) {
    if (!path || !out) return set_error(PT_ERR_INVALID, "pt_scene_file_load: null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(PT_ERR_INVALID, "cannot open %s", path);
    pt_scene_file* s = new (std::nothrow) pt_scene_file();
    if (!s) {
        std::fclose(f);
        return set_error(PT_ERR_MEMORY, "out of host memory");
    }
    auto bad = [&](const char* why) {
        std::fclose(f);
        delete s;
        return set_error(PT_ERR_INVALID, "%s: %s", path, why);
    };
    char magic[8];
    uint32_t count = 0, pad = 0;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "PTSCENE1", 8) != 0) return bad("not a .ptscene file");
    if (std::fread(&count, 4, 1, f) != 1 || std::fread(&pad, 4, 1, f) != 1 || count > 64) return bad("bad header");
    std::fseek(f, 0, SEEK_END);
    const long file_size = std::ftell(f);
    std::fseek(f, 16, SEEK_SET);
    try {
        for (uint32_t i = 0; i < count; ++i) {
            uint32_t hdr[2];
            uint64_t n = 0;
            if (std::fread(hdr, 4, 2, f) != 2 || std::fread(&n, 8, 1, f) != 1) return bad("truncated section header");
            const long at = std::ftell(f);
            if (at < 0 || n > (uint64_t)(file_size - at)) return bad("section larger than the file");
            // each section is read straight into its array (no second copy of large ones)
            auto take = [&](auto& vec) {
                using T = typename std::decay_t<decltype(vec)>::value_type;
                if (n % sizeof(T) || n / sizeof(T) > 0xffffffffull) return false;
                vec.resize(n / sizeof(T));
                return n == 0 || std::fread(vec.data(), 1, n, f) == n;
            };
            auto take_one = [&](void* dst, size_t size) { return n == size && std::fread(dst, 1, size, f) == size; };
            bool ok = true;
            switch (hdr[0]) {
                case tag("GEOM"): ok = take(s->geo); break;
                case tag("ATTR"): ok = take(s->attr); break;
                case tag("MATL"): ok = take(s->mat); break;
                case tag("BVHN"): ok = take(s->bvh); break;
                case tag("ASIZ"): ok = take_one(&s->atlas_size, 4); break;
                case tag("SRGB"): ok = take(s->srgb); break;
                case tag("LINR"): ok = take(s->linear); break;
                case tag("CAMR"): ok = take_one(&s->camera, sizeof(PtCamera)); break;
                case tag("SKYU"): ok = take_one(&s->skybox, sizeof(PtSkyboxUniforms)); break;
                case tag("SKYR"): ok = take(s->sky); break;
                case tag("SKYA"): ok = take(s->sky_alias); break;
                case tag("SCNU"): ok = take_one(&s->uniforms, sizeof(PtSceneUniforms)); break;
                case tag("SCNA"): ok = take(s->scene_alias); break;
                default: ok = std::fseek(f, (long)n, SEEK_CUR) == 0; break;  // unknown sections are skipped
            }
            if (!ok) return bad("section size does not fit its records, or the section is truncated");
            if (std::fseek(f, (long)((16 - n % 16) % 16), SEEK_CUR) != 0) return bad("truncated section padding");
        }
    } catch (const std::bad_alloc&) {
        std::fclose(f);
        delete s;
        return set_error(PT_ERR_MEMORY, "out of host memory reading %s", path);
    }
    std::fclose(f);
    const size_t layer = (size_t)s->atlas_size * s->atlas_size * 4;
    if (s->attr.size() != s->geo.size()) return delete s, set_error(PT_ERR_INVALID, "%s: attribute count differs", path);
    if (layer ? (s->srgb.size() % layer || s->linear.size() % layer) : (!s->srgb.empty() || !s->linear.empty()))
        return delete s, set_error(PT_ERR_INVALID, "%s: atlas bytes are not whole layers", path);
    s->srgb_count = layer ? (uint32_t)(s->srgb.size() / layer) : 0;
    s->linear_count = layer ? (uint32_t)(s->linear.size() / layer) : 0;
    if (s->sky.size() != (size_t)s->skybox.width * s->skybox.height * 4)
        return delete s, set_error(PT_ERR_INVALID, "%s: sky size differs from its uniforms", path);
    if (s->srgb_count > 0x7fffffffu || s->linear_count > 0x7fffffffu)
        return delete s, set_error(PT_ERR_INVALID, "%s: too many atlas layers", path);
    // the indices the kernels follow, checked as pt_scene_create checks them
    pt_scene_desc d;
    uint32_t depth = 0;
    pt_scene_file_desc(s, &d);
    if (pt_status st = validate_scene(&d, depth)) {
        const std::string why = pt_last_error();
        delete s;
        return set_error(st, "%s: %s", path, why.c_str());
    }
    *out = s;
    return PT_OK;
}

pt_status pt_scene_file_desc(const pt_scene_file* s, pt_scene_desc* d) {
    if (!s || !d) return set_error(PT_ERR_INVALID, "null argument");
    std::memset(d, 0, sizeof(*d));
    d->geometries = s->geo.data();
    d->attributes = s->attr.data();
    d->triangle_count = (uint32_t)s->geo.size();
    d->materials = s->mat.data();
    d->material_count = (uint32_t)s->mat.size();
    d->bvh_nodes = s->bvh.data();
    d->bvh_node_count = (uint32_t)s->bvh.size();
    d->srgb_atlases = s->srgb.empty() ? nullptr : s->srgb.data();
    d->srgb_atlas_count = s->srgb_count;
    d->linear_atlases = s->linear.empty() ? nullptr : s->linear.data();
    d->linear_atlas_count = s->linear_count;
    d->atlas_size = s->atlas_size;
    d->camera = s->camera;
    d->skybox_rgba = s->sky.data();
    d->skybox = s->skybox;
    d->skybox_alias = s->sky_alias.empty() ? nullptr : s->sky_alias.data();
    d->skybox_alias_count = (uint32_t)s->sky_alias.size();
    d->scene_uniforms = s->uniforms;
    d->scene_alias = s->scene_alias.empty() ? nullptr : s->scene_alias.data();
    d->scene_alias_count = (uint32_t)s->scene_alias.size();
    return PT_OK;
}

pt_status pt_scene_file_destroy(pt_scene_file* s) {
    delete s;
    return PT_OK;
}

}  // extern "C"
