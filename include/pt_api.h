/*
 * pt_api.h - C ABI of libptcore.so, the MI355X path-tracing core.
 *
 * This is the boundary a host binds (ctypes, a Rust FFI crate, cgo ...). It
 * replaces the reference's WebGPU compute pipeline for trace.slang and the
 * Rust objects that drive it:
 *
 *   pt_context_*   <- gpu::request_adapter/request_device     src/render/gpu.rs:14-76
 *   pt_world_*     <- World::new + World::take_updates (BVH build, triangle
 *                     reorder, alias tables, sky normalisation, camera)
 *                                                              src/render/world.rs:589-884, 1339-1355,
 *                                                              1430-1591, 1691-1741, 1872-1987, 2159-2193
 *   pt_scene_*     <- WorldResources::{new, apply_updates} (bind group 1)
 *                                                              src/render/world_resources.rs:249-390, 479-525, 600-661
 *   pt_tracer_*    <- PathTracer::{new, set_tile, set_max_bounces,
 *                     apply_updates, trace, display, sample_count}
 *                                                              src/render/path_tracer.rs:59-358
 *   pt_batch_*, pt_server_* : batching of many scenes and a serving loop;
 *                     no counterpart in the reference (it renders one tile
 *                     of one scene at a time, headless.rs:207-285).
 *
 * Conventions: every call returns pt_status (PT_OK = 0); on failure
 * pt_last_error() returns a thread-local message. Host arrays passed in are
 * borrowed for the duration of the call only (the library copies them).
 * Objects are not thread-safe; different contexts may be used from different
 * threads. All structs are the byte layouts of pt_layout.h.
 */
#ifndef PT_API_H
#define PT_API_H

#include <stddef.h>
#include <stdint.h>

#include "pt_layout.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pt_status;
#define PT_OK 0
#define PT_ERR_INVALID 1   /* bad argument, like the reference's anyhow errors on limits */
#define PT_ERR_DEVICE 2    /* HIP runtime error */
#define PT_ERR_MEMORY 3    /* allocation failed or over the admission budget */
#define PT_ERR_STATE 4     /* call not valid in the current state */
#define PT_ERR_TIMEOUT 5   /* a job did not finish in time */
#define PT_ERR_CANCELLED 6 /* a job was cancelled */

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* pt_last_error(void);
/* Library version string. */
const char* pt_version(void);

/* ------------------------------------------------------------------ context */
typedef struct pt_context pt_context;

/* Select HIP device `device` and create a stream for this context.
 * Replaces gpu.rs:14-76 (adapter + device request). */
pt_status pt_context_create(int device, pt_context** out);
pt_status pt_context_destroy(pt_context* ctx);
/* Wait for all work queued on the context's stream. */
pt_status pt_context_synchronize(pt_context* ctx);

/* -------------------------------------------------------------- host world */
/* Input of the host scene preparation: a triangle soup in input order (the
 * order glTF flattening produces, world.rs:1114-1337), materials, atlases,
 * an equirectangular RGBA32F sky and a camera. */
typedef struct pt_world_desc {
    const PtTriangleGeometry* geometries; /* triangle_count, input order */
    const PtTriangleAttribute* attributes;
    uint32_t triangle_count;
    const PtMaterial* materials;
    uint32_t material_count;
    const uint8_t* srgb_atlases;   /* srgb_atlas_count layers of atlas_size^2 RGBA8 */
    uint32_t srgb_atlas_count;
    const uint8_t* linear_atlases; /* linear_atlas_count layers */
    uint32_t linear_atlas_count;
    uint32_t atlas_size;
    const float* skybox_rgba;      /* skybox_width * skybox_height * 4, un-normalised */
    uint32_t skybox_width;
    uint32_t skybox_height;
    float skybox_luminance;        /* average luminance in nits; < 0 keeps the image's own */
    /* camera: NaN in a field means "autogenerated" (Camera::new, world.rs:2178-2193) */
    float camera_position[3];
    float camera_yaw_degrees;
    float camera_pitch_degrees;
    float camera_fov_degrees;
    float aspect_ratio;            /* width / height */
} pt_world_desc;

typedef struct pt_world pt_world;

/* Build the BVH (binned SAH, 8 bins), reorder triangles into BVH order, build
 * the emissive-triangle and sky alias tables, normalise the sky and compute
 * the camera, exactly as World::new + World::take_updates do. */
pt_status pt_world_build(const pt_world_desc* desc, pt_world** out);
pt_status pt_world_destroy(pt_world* world);

/* Device-layout arrays of a scene (bind group 1 of the reference). */
typedef struct pt_scene_desc {
    const PtTriangleGeometry* geometries;
    const PtTriangleAttribute* attributes;
    uint32_t triangle_count;
    const PtMaterial* materials;
    uint32_t material_count;
    const PtBvhNode* bvh_nodes;
    uint32_t bvh_node_count;
    const uint8_t* srgb_atlases;
    uint32_t srgb_atlas_count;
    const uint8_t* linear_atlases;
    uint32_t linear_atlas_count;
    uint32_t atlas_size;
    PtCamera camera;
    const float* skybox_rgba;                 /* normalised, width*height*4 */
    PtSkyboxUniforms skybox;
    const PtSkyboxAliasEntry* skybox_alias;   /* width*height entries, or 0 entries for a black sky */
    uint32_t skybox_alias_count;
    PtSceneUniforms scene_uniforms;
    const PtSceneAliasEntry* scene_alias;     /* scene_uniforms.emissive_triangles_count entries */
    uint32_t scene_alias_count;
} pt_scene_desc;

/* Borrow the prepared arrays of `world` (valid until pt_world_destroy). */
pt_status pt_world_scene_desc(const pt_world* world, pt_scene_desc* out);

/* ------------------------------------------------------------ device scene */
typedef struct pt_scene pt_scene;

/* Copy a scene to device memory owned by the library (WorldResources::new). */
pt_status pt_scene_create(pt_context* ctx, const pt_scene_desc* desc, pt_scene** out);
pt_status pt_scene_destroy(pt_scene* scene);
/* Device bytes held by the scene (used for admission by HBM footprint). */
pt_status pt_scene_footprint_bytes(const pt_scene* scene, uint64_t* bytes);

/* Partial updates (WorldResources::apply_updates, world_resources.rs:479-525).
 * Each one bumps the scene's version so tracers reset their accumulation
 * (path_tracer.rs:272-282). Arrays are borrowed for the call. */
pt_status pt_scene_update_camera(pt_scene* scene, const PtCamera* camera);
pt_status pt_scene_update_materials(pt_scene* scene, const uint32_t* indices,
                                    const PtMaterial* materials, uint32_t count);
pt_status pt_scene_update_lights(pt_scene* scene, const PtSceneUniforms* uniforms,
                                 const PtSceneAliasEntry* alias, uint32_t alias_count,
                                 const PtSkyboxUniforms* skybox);

/* Closest-hit queries on the device BVH for arbitrary rays (host arrays).
 * rays: count * 6 floats (origin xyz, direction xyz). Outputs per ray: hit
 * triangle index (-1 on miss), t, barycentrics v, w. Optional counters (may be
 * NULL): child-pair box tests and triangle tests per ray. */
pt_status pt_scene_intersect(pt_scene* scene, const float* rays, uint32_t count,
                             int32_t* hit_index, float* hit_t, float* hit_v, float* hit_w,
                             uint32_t* pair_tests, uint32_t* tri_tests);
/* Any-hit (shadow) queries: occluded[i] = 1 if a surface lies in (0, tmax[i]),
 * skipping triangle skip_index[i] (-1 for none). */
pt_status pt_scene_occluded(pt_scene* scene, const float* rays, const float* tmax,
                            const int32_t* skip_index, uint32_t count, uint8_t* occluded);

/* ------------------------------------------------------------------ tracer */
typedef struct pt_tracer pt_tracer;

/* Kernel schedules. MEGAKERNEL follows the reference's one-thread-per-path
 * structure; WAVEFRONT splits each bounce into generate/extend/shade/shadow
 * kernels over compacted queues. Both give the same per-pixel result. */
#define PT_SCHEDULE_MEGAKERNEL 0
#define PT_SCHEDULE_WAVEFRONT 1

/* PathTracer::new: a tracer over `scene` with a tile of tile_w x tile_h. */
pt_status pt_tracer_create(pt_context* ctx, pt_scene* scene, uint32_t tile_w, uint32_t tile_h,
                           pt_tracer** out);
pt_status pt_tracer_destroy(pt_tracer* tracer);
pt_status pt_tracer_set_schedule(pt_tracer* tracer, int schedule);
/* PathTracer::set_tile: resets accumulation; the origin must be inside the image. */
pt_status pt_tracer_set_tile(pt_tracer* tracer, uint32_t image_w, uint32_t image_h,
                             uint32_t origin_x, uint32_t origin_y, uint32_t tile_w, uint32_t tile_h);
/* PathTracer::set_max_bounces: resets accumulation when it changes; 0..16. */
pt_status pt_tracer_set_max_bounces(pt_tracer* tracer, uint32_t max_bounces);
/* Restore accumulation state (resume): the next sample traced is
 * sample_count + 1. The accumulation buffer is left as it is. */
pt_status pt_tracer_set_sample_count(pt_tracer* tracer, uint32_t sample_count);
uint32_t pt_tracer_sample_count(const pt_tracer* tracer);
uint32_t pt_tracer_max_bounces(const pt_tracer* tracer);
/* PathTracer::trace: queue up to min(samples, 16) samples of the tile on the
 * context's stream; *traced receives the number queued. Asynchronous. */
pt_status pt_tracer_trace(pt_tracer* tracer, uint32_t samples, uint32_t* traced);
/* Queue `samples` samples with no per-call cap (one wavefront launch
 * sequence over all of them). Asynchronous. */
pt_status pt_tracer_trace_all(pt_tracer* tracer, uint32_t samples);
/* Copy the accumulation buffer (tile_w*tile_h float4, sum of samples) or
 * the storage texture (EXPOSURE * sum / sample_count) to host memory.
 * Synchronises the stream. */
pt_status pt_tracer_read_accumulation(pt_tracer* tracer, float* rgba);
pt_status pt_tracer_write_accumulation(pt_tracer* tracer, const float* rgba);
pt_status pt_tracer_read_storage(pt_tracer* tracer, float* rgba);
/* PathTracer::display + readback: ACES filmic tone map and sRGB encode of the
 * storage texture into RGBA8 (display.wgsl:19-46). */
pt_status pt_tracer_display_rgba8(pt_tracer* tracer, uint8_t* rgba);
/* Primary-ray closest hits of the tile for 1-based sample index
 * `sample_count` (same jitter as the trace): triangle index (-1 on miss), t. */
pt_status pt_tracer_primary_hits(pt_tracer* tracer, uint32_t sample_count, int32_t* tri, float* t);

/* Device counters of the last trace call (wavefront schedule): closest-hit
 * rays, shadow rays, child-pair box tests, triangle tests, path segments. */
typedef struct pt_trace_stats {
    uint64_t closest_rays;
    uint64_t shadow_rays;
    uint64_t pair_tests;
    uint64_t tri_tests;
    uint64_t extend_launches;
    float extend_ms;      /* summed HIP-event time of the extend kernels */
    float shadow_ms;      /* summed HIP-event time of the shadow kernels */
    float shade_ms;
    float total_ms;
} pt_trace_stats;
/* Enable/disable device counters and per-kernel event timing (off by default;
 * on costs a few percent). */
pt_status pt_tracer_set_profiling(pt_tracer* tracer, int enabled);
pt_status pt_tracer_stats(pt_tracer* tracer, pt_trace_stats* out);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* PT_API_H */
