// pt_internal.h - host-side internals shared by the library's translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/pt_api.h"

namespace ptcore {

// Thread-local error message for pt_last_error(); returns `code`.
pt_status set_error(pt_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// `s` with every byte that is not part of well-formed UTF-8 text replaced by '?'.
std::string utf8_clean(const char* s);
// The checks pt_scene_create makes before upload: every index the kernels
// follow is in range and the BVH is a tree (its depth is returned).
pt_status validate_scene(const pt_scene_desc* d, uint32_t& depth);

// Host-side prepared world (the arrays World::take_updates hands to the GPU).
struct World {
    std::vector<PtTriangleGeometry> geometries;
    std::vector<PtTriangleAttribute> attributes;
    std::vector<PtMaterial> materials;
    std::vector<PtBvhNode> bvh_nodes;
    std::vector<uint8_t> srgb, linear;
    uint32_t srgb_count = 0, linear_count = 0, atlas_size = 0;
    std::vector<float> sky;
    PtSkyboxUniforms skybox{};
    std::vector<PtSkyboxAliasEntry> sky_alias;
    PtSceneUniforms scene_uniforms{};
    std::vector<PtSceneAliasEntry> scene_alias;
    PtCamera camera{};
    // editing state (World, world.rs:53-71 and 589-884)
    pt_camera_params cam{}, initial_cam{};
    std::vector<PtMaterial> initial_materials;
    double radius_squared = 0.0;
    float emission_scale = 1.0f;
    bool pending_camera = false, pending_luminance = false;
    pt_camera_params pending_cam{};
    float pending_lum = 0.0f;
    std::map<uint32_t, PtMaterial> pending_materials;  // ascending index, like the reference's BTreeMap
    std::vector<uint32_t> update_indices;              // arrays of the last take_updates
    std::vector<PtMaterial> update_materials;
    int32_t selected_material = -1;
};

std::vector<PtSkyboxAliasEntry> alias_table(const std::vector<double>& weights);
void update_light_probabilities(PtSceneUniforms& u, float skybox_flux);
PtCamera gpu_camera(const float pos[3], float yaw, float pitch, float fov, float aspect);
int32_t pick_material(const World& w, const float pos[3], float yaw, float pitch);

}  // namespace ptcore

struct pt_world : ptcore::World {};
