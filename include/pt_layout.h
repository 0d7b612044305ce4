/*
 * pt_layout.h - device byte layouts of the path-tracing core.
 *
 * Every struct here is byte-for-byte the layout the reference uploads to its
 * GPU buffers, so a scene serialized by the reference's host (or by ours)
 * is consumed unchanged by the HIP kernels, the C++ oracle and the host
 * prep code. Sizes and offsets are pinned with static_asserts.
 *
 *   PtTriangleGeometry  <- GpuTriangleGeometry   src/render/world.rs:139-146
 *   PtTriangleAttribute <- GpuTriangleAttribute  src/render/world.rs:155-182
 *   PtMaterial          <- GpuMaterial           src/render/world.rs:199-309
 *   PtBvhNode           <- GpuBvhNode            src/render/world.rs:323-335
 *   PtCamera            <- GpuCamera             src/render/world.rs:389-398
 *   PtSceneUniforms     <- GpuSceneUniforms      src/render/world.rs:476-482
 *   PtSceneAliasEntry   <- GpuSceneAliasEntry    src/render/world.rs:503-510
 *   PtSkyboxUniforms    <- GpuSkyboxUniforms     src/render/world.rs:522-529
 *   PtSkyboxAliasEntry  <- GpuSkyboxAliasEntry   src/render/world.rs:533-537
 *   PtTraceParameters   <- GpuTraceParameters    src/render/path_tracer.rs:45-57
 *
 * Plain C so that ctypes, a Rust/Go FFI binding and the HIP kernels can all
 * include it.
 */
#ifndef PT_LAYOUT_H
#define PT_LAYOUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Alpha modes (src/render/world.rs:31-35, trace.slang:109-110). */
#define PT_ALPHA_MODE_OPAQUE 0u
#define PT_ALPHA_MODE_MASK 1u
#define PT_ALPHA_MODE_BLEND 2u

/* Constants of the kernel and driver. */
#define PT_EXPOSURE 0.00002604166f          /* trace.slang:25 */
#define PT_DEFAULT_MAX_BOUNCES 10u          /* path_tracer.rs:61 */
#define PT_MAX_BOUNCES_LIMIT 16u            /* path_tracer.rs:63 */
#define PT_MAX_SAMPLES_PER_TRACE 16u        /* path_tracer.rs:66 */
#define PT_ATLAS_SIZE 8192u                 /* world.rs:29 */
#define PT_BVH_BINS 8                       /* world.rs:437 */
#define PT_DEFAULT_SKYBOX_LUMINANCE 8000.0f /* world.rs:21 */

/* Three vertices, x,y,z each, packed without padding (36 B). */
typedef struct PtTriangleGeometry {
    float v[9];
} PtTriangleGeometry;

/* Material index and per-vertex normals, UVs and tangents (112 B, 16-aligned). */
typedef struct PtTriangleAttribute {
    uint32_t material;  /* index into the material array */
    float n[9];         /* n0, n1, n2 */
    float uv[6];        /* uv0, uv1, uv2 */
    float t0[4];        /* tangent + handedness in w */
    float t1[4];
    float t2[4];
} PtTriangleAttribute;

/* OpenPBR material (224 B). Field order follows the reference exactly. */
typedef struct PtMaterial {
    float base_color[4];
    float base_weight;
    float base_metalness;
    float base_diffuse_roughness;
    float emission_weight;
    float emission_color_and_luminance[4];
    float specular_color_and_weight[4];
    float specular_roughness;
    float specular_roughness_anisotropy;
    float specular_haze;
    float specular_haze_spread;
    float specular_ior;
    float specular_retroreflectivity;
    float normal_scale;
    uint32_t alpha_mode;
    float alpha_cutoff;
    float pad;
    int32_t base_color_tex_layer;
    int32_t normal_tex_layer;
    int32_t metallic_roughness_tex_layer;
    int32_t emission_tex_layer;
    int32_t specular_color_tex_layer;
    int32_t specular_weight_tex_layer;
    float base_color_uv[4];       /* offset.xy, scale.xy inside the atlas */
    float normal_uv[4];
    float metallic_roughness_uv[4];
    float emission_uv[4];
    float specular_color_uv[4];
    float specular_weight_uv[4];
} PtMaterial;

/* BVH node (32 B). Internal: tri_count == 0 and the children are the adjacent
 * pair (left_first, left_first + 1), i.e. one 64-byte pair in memory.
 * Leaf: triangles [left_first, left_first + tri_count). */
typedef struct PtBvhNode {
    float aabb_min[3];
    uint32_t left_first;
    float aabb_max[3];
    uint32_t tri_count;
} PtBvhNode;

/* Pinhole camera (64 B); the 4th component of each vector is unused. */
typedef struct PtCamera {
    float position[4];
    float lower_left_corner_minus_position[4];
    float horizontal[4];
    float vertical[4];
} PtCamera;

typedef struct PtSceneUniforms {
    float total_luminous_flux;
    uint32_t emissive_triangles_count;
    float light_probabilities[2]; /* geometry, skybox */
} PtSceneUniforms;

typedef struct PtSceneAliasEntry {
    float threshold;
    uint32_t alias_index;
    uint32_t triangle_index;
    float selection_probability;
} PtSceneAliasEntry;

typedef struct PtSkyboxUniforms {
    float total_luminous_flux;
    float average_luminance;
    uint32_t width;
    uint32_t height;
} PtSkyboxUniforms;

typedef struct PtSkyboxAliasEntry {
    float threshold;
    uint32_t alias_index;
    float selection_probability;
} PtSkyboxAliasEntry;

/* Per-sample parameters (32 B). sample_count is 1-based and counts the
 * samples accumulated in the tile including the one being traced. */
typedef struct PtTraceParameters {
    uint32_t image_size[2];
    uint32_t tile_origin[2];
    uint32_t max_bounces;
    uint32_t sample_count;
    uint32_t pad[2];
} PtTraceParameters;

#ifdef __cplusplus
}  /* extern "C" */
#endif

#if defined(__cplusplus)
#include <cstddef>
#define PT_LAYOUT_ASSERT(c, m) static_assert(c, m)
#else
#include <stddef.h>
#define PT_LAYOUT_ASSERT(c, m) _Static_assert(c, m)
#endif

PT_LAYOUT_ASSERT(sizeof(PtTriangleGeometry) == 36, "geometry size");
PT_LAYOUT_ASSERT(sizeof(PtTriangleAttribute) == 112, "attribute size");
PT_LAYOUT_ASSERT(offsetof(PtTriangleAttribute, n) == 4, "attribute n0");
PT_LAYOUT_ASSERT(offsetof(PtTriangleAttribute, uv) == 40, "attribute uv0");
PT_LAYOUT_ASSERT(offsetof(PtTriangleAttribute, t0) == 64, "attribute t0");
PT_LAYOUT_ASSERT(offsetof(PtTriangleAttribute, t1) == 80, "attribute t1");
PT_LAYOUT_ASSERT(offsetof(PtTriangleAttribute, t2) == 96, "attribute t2");
PT_LAYOUT_ASSERT(sizeof(PtMaterial) == 224, "material size");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, base_weight) == 16, "material base_weight");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, emission_color_and_luminance) == 32, "material emission");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, specular_color_and_weight) == 48, "material specular");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, specular_roughness) == 64, "material spec roughness");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, alpha_mode) == 92, "material alpha_mode");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, base_color_tex_layer) == 104, "material tex layers");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, base_color_uv) == 128, "material uv");
PT_LAYOUT_ASSERT(offsetof(PtMaterial, specular_weight_uv) == 208, "material last uv");
PT_LAYOUT_ASSERT(sizeof(PtBvhNode) == 32, "bvh node size");
PT_LAYOUT_ASSERT(offsetof(PtBvhNode, left_first) == 12, "bvh left_first");
PT_LAYOUT_ASSERT(offsetof(PtBvhNode, tri_count) == 28, "bvh tri_count");
PT_LAYOUT_ASSERT(sizeof(PtCamera) == 64, "camera size");
PT_LAYOUT_ASSERT(sizeof(PtSceneUniforms) == 16, "scene uniforms size");
PT_LAYOUT_ASSERT(sizeof(PtSceneAliasEntry) == 16, "scene alias size");
PT_LAYOUT_ASSERT(sizeof(PtSkyboxUniforms) == 16, "skybox uniforms size");
PT_LAYOUT_ASSERT(sizeof(PtSkyboxAliasEntry) == 12, "skybox alias size");
PT_LAYOUT_ASSERT(sizeof(PtTraceParameters) == 32, "trace parameters size");

#endif /* PT_LAYOUT_H */
