// pt_internal.h - host-side internals shared by the library's translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <vector>

#include "../../include/pt_api.h"

namespace ptcore {

// Thread-local error message for pt_last_error(); returns `code`.
pt_status set_error(pt_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

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
};

std::vector<PtSkyboxAliasEntry> alias_table(const std::vector<double>& weights);
void update_light_probabilities(PtSceneUniforms& u, float skybox_flux);
PtCamera gpu_camera(const float pos[3], float yaw, float pitch, float fov, float aspect);

}  // namespace ptcore

struct pt_world : ptcore::World {};
