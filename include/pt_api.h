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

/* Select HIP device `device` and create the context's two streams: all work
 * of the context runs on them (a render's second lane and a render server's
 * second lane on the second one).
 * Replaces gpu.rs:14-76 (adapter + device request). */
pt_status pt_context_create(int device, pt_context** out);
/* Waits for both streams, then releases the context. Servers, tracers and
 * scenes made on it must be destroyed first. */
pt_status pt_context_destroy(pt_context* ctx);
/* Wait for all work queued on the context's two streams. */
pt_status pt_context_synchronize(pt_context* ctx);
/* The device facts behind launch sizes and rate reports; replaces the
 * adapter.limits() / adapter.get_info() queries of gpu.rs:39-44. */
typedef struct pt_context_info {
    int32_t device;
    uint32_t cu_count;      /* compute units */
    uint32_t clock_khz;     /* peak engine clock */
    uint32_t l2_bytes;      /* L2 cache of one XCD */
    uint64_t hbm_bytes;     /* device memory */
    char arch[64];          /* e.g. "gfx950:sramecc+:xnack-" */
} pt_context_info;
pt_status pt_context_get_info(const pt_context* ctx, pt_context_info* info);
/* Measures the vector L1's gather rate on the context's GPU: lane accesses
 * per CU per cycle (at the peak clock) of 16-byte loads from an L1-resident
 * buffer, every lane at a pseudo-random 64-byte segment (lanes_per_segment 1:
 * the traversal's divergent child-pair loads) or four consecutive lanes at the
 * four parts of one segment (lanes_per_segment 4). Times a probe kernel with
 * HIP events on the context's stream and synchronises it. No counterpart in
 * the reference: it is the measured ceiling that rate reports set the
 * traversal's lane-access rate against. */
pt_status pt_context_measure_gather(pt_context* ctx, uint32_t lanes_per_segment, double* lane_accesses_per_cu_cycle);
/* Checks, on the context's GPU, the shading code's restricted-domain square
 * root and division against the device's correctly rounded sqrtf and
 * division. The root is compared on every float32 bit pattern from
 * sqrt_first to sqrt_last (inclusive); the division on div_pairs pseudo-random
 * operand pairs from its domain (divisor magnitude in [2^-60, 2^60), dividend
 * +0 or of magnitude in [2^-60, 2^4), both signs), drawn from div_seed.
 * Results that differ (NaN equals NaN) are counted into *sqrt_mismatches and
 * *div_mismatches. Synchronises the context's stream. No counterpart in the
 * reference: it pins the arithmetic that keeps the kernels bit-exact. */
pt_status pt_context_check_fast_math(pt_context* ctx, uint32_t sqrt_first, uint32_t sqrt_last, uint64_t div_pairs,
                                     uint32_t div_seed, uint64_t* sqrt_mismatches, uint64_t* div_mismatches);

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

/* Borrow the prepared arrays of `world` (valid until pt_world_destroy or the
 * next pt_world_take_updates, whose edits may replace the alias table). */
pt_status pt_world_scene_desc(const pt_world* world, pt_scene_desc* out);

/* ------------------------------------------------------- world editing */
/* Camera parameters (world.rs:360-376): angles in radians, pitch in
 * [-pi/2, pi/2], vertical fov in [0, pi). */
typedef struct pt_camera_params {
    float position[3];
    float yaw;
    float pitch;
    float fov;
    float aspect_ratio;
} pt_camera_params;

/* Current (initial == 0) or initial camera (World::camera / initial_camera,
 * world.rs:700-706). Edits made since the last take_updates are not included. */
pt_status pt_world_camera(const pt_world* world, int initial, pt_camera_params* out);
/* World::edit_camera (world.rs:708-716): replaces the pending camera. */
pt_status pt_world_edit_camera(pt_world* world, const pt_camera_params* camera);
/* Current and initial material `index` (World::material, world.rs:725-734);
 * either output may be null. */
pt_status pt_world_material(const pt_world* world, uint32_t index, PtMaterial* current, PtMaterial* initial);
/* World::edit_material (world.rs:736-783): the fields before alpha_mode are
 * taken from `candidate` when all of them are finite (*accepted = 1), else
 * the edit is ignored (*accepted = 0); alpha mode, cutoff and texture
 * bindings stay as they are. */
pt_status pt_world_edit_material(pt_world* world, uint32_t index, const PtMaterial* candidate, int* accepted);
/* World::edit_emission_scale (world.rs:654-690): every emissive material's
 * emission luminance becomes its initial one times max(scale, 0). */
pt_status pt_world_edit_emission_scale(pt_world* world, float scale);
pt_status pt_world_emission_scale(const pt_world* world, float* scale);
/* World::edit_skybox_average_luminance (world.rs:644-648), in nits (>= 0). */
pt_status pt_world_edit_skybox_luminance(pt_world* world, float luminance);
/* Scene::pick_material (world.rs:1595-1688): the material of the triangle a
 * ray from `position` along the camera's forward axis hits first, or -1. */
pt_status pt_world_pick_material(const pt_world* world, const float position[3], float yaw, float pitch,
                                 int32_t* material);

/* What changed since the last take (WorldUpdates, world.rs:73-101). Arrays
 * point into the world and stay valid until its next take or destroy. */
typedef struct pt_world_updates {
    uint32_t camera_changed;
    PtCamera camera;
    uint32_t material_count;               /* edited materials, ascending index */
    const uint32_t* material_indices;
    const PtMaterial* materials;
    uint32_t scene_uniforms_changed;
    PtSceneUniforms scene_uniforms;
    uint32_t scene_alias_changed;          /* emissive set or emission changed */
    const PtSceneAliasEntry* scene_alias;
    uint32_t scene_alias_count;
    uint32_t skybox_uniforms_changed;
    PtSkyboxUniforms skybox_uniforms;
    uint32_t reset_accumulation;
    int32_t selected_material;             /* pick_material at the camera, -1 for none */
} pt_world_updates;
/* World::take_updates (world.rs:785-884): applies the pending edits to the
 * world, rebuilding the emissive alias table and the light probabilities when
 * emission or sky luminance changed, and reports what a device scene needs. */
pt_status pt_world_take_updates(pt_world* world, pt_world_updates* out);

/* ------------------------------------------------------------ device scene */
typedef struct pt_scene pt_scene;

/* Copy a scene to device memory owned by the library (WorldResources::new). */
pt_status pt_scene_create(pt_context* ctx, const pt_scene_desc* desc, pt_scene** out);
/* PT_ERR_STATE (scene kept) while a render server has jobs over the scene
 * that are not done, failed or cancelled. */
pt_status pt_scene_destroy(pt_scene* scene);
/* Device bytes held by the scene (used for admission by HBM footprint). */
pt_status pt_scene_footprint_bytes(const pt_scene* scene, uint64_t* bytes);

/* Partial updates (WorldResources::apply_updates, world_resources.rs:479-525).
 * Each one bumps the scene's version so tracers reset their accumulation
 * (path_tracer.rs:272-282). Arrays are borrowed for the call. While server
 * jobs over the scene are unfinished, they return PT_ERR_STATE and change
 * nothing, so a job never mixes two versions of a scene. */
pt_status pt_scene_update_camera(pt_scene* scene, const PtCamera* camera);
pt_status pt_scene_update_materials(pt_scene* scene, const uint32_t* indices,
                                    const PtMaterial* materials, uint32_t count);
pt_status pt_scene_update_lights(pt_scene* scene, const PtSceneUniforms* uniforms,
                                 const PtSceneAliasEntry* alias, uint32_t alias_count,
                                 const PtSkyboxUniforms* skybox);
/* PathTracer::apply_updates (path_tracer.rs:272-282) for what
 * pt_world_take_updates reported: writes the changed parts to the device and
 * resets tracers' accumulation only when reset_accumulation is set. */
pt_status pt_scene_apply_updates(pt_scene* scene, const pt_world_updates* updates);

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
/* No counterpart in the reference: a tile render for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It traces `samples` consecutive samples,
 * from the 1-based index params->sample_count, of the tile of params'
 * image at params->tile_origin, with the megakernel schedule (the same
 * kernel, and so the same bits, as pt_tracer_trace_all on a megakernel
 * tracer with that tile). `accum` is the caller's device buffer of
 * tile_w * tile_h float4, read (when the first sample is > 1) and written as
 * the running sum; `storage` (device, may be NULL) receives EXPOSURE * sum /
 * last sample index. The kernel's traversal overflow stacks use a static
 * device array of 32 MiB per GPU (held until the process ends); the call
 * allocates and frees nothing, and rejects a scene whose BVH is too deep for
 * it (PT_ERR_INVALID). Every call on one GPU shares that array: calls that
 * overlap on one GPU (other contexts, other threads) race. Asynchronous on the
 * context's stream (pt_context_synchronize orders it with the caller's use of
 * the buffers). */
pt_status pt_render_tile_noalloc(pt_scene* scene, const PtTraceParameters* params, uint32_t tile_w, uint32_t tile_h,
                                 uint32_t samples, float* accum, float* storage);
/* The display pass on host data (display.wgsl:19-46): the ACES fit and the
 * sRGB encode of `count` float4 texels (rgba, count * 4 floats) into RGBA8,
 * rounded to nearest as an Rgba8Unorm target stores them (alpha clamped to
 * [0, 1]). The same kernel as pt_tracer_display_rgba8, on the context's stream;
 * synchronises. */
pt_status pt_tonemap_rgba8(pt_context* ctx, const float* rgba, uint32_t count, uint8_t* out);
/* Primary-ray closest hits of the tile for 1-based sample index
 * `sample_count` (same jitter as the trace): triangle index (-1 on miss), t. */
pt_status pt_tracer_primary_hits(pt_tracer* tracer, uint32_t sample_count, int32_t* tri, float* t);

/* Profiling of the trace calls made since pt_tracer_set_profiling: HIP-event
 * time per kernel class (recorded on the context's stream around each
 * launch, read without per-launch synchronisation) and, at level 2, device
 * traversal counters: closest-hit and shadow rays issued, child-pair box
 * tests (64 contiguous bytes each) and triangle tests (36 bytes each). The
 * megakernel reports its one kernel in total_ms and all traversals in
 * pair_tests/tri_tests. */
typedef struct pt_trace_stats {
    uint64_t closest_rays;
    uint64_t shadow_rays;
    uint64_t pair_tests;         /* closest-hit traversal */
    uint64_t tri_tests;
    uint64_t shadow_pair_tests;  /* any-hit traversal (wavefront) */
    uint64_t shadow_tri_tests;
    uint64_t extend_launches;
    float extend_ms;             /* summed HIP-event time of the extend kernels */
    float shadow_ms;             /* summed HIP-event time of the shadow kernels */
    float shade_ms;
    float total_ms;              /* whole trace calls */
    float extend_busy_ms;        /* time at least one extend kernel was running: the union of their
                                    event intervals (launches on two lanes may overlap, so this is at
                                    most the elapsed time, while extend_ms may exceed it) */
} pt_trace_stats;
/* level 0 off, 1 per-kernel event timing, 2 timing + device counters
 * (counting kernels; synchronises once per chunk), 3 event timing of the
 * traversal launches only (extend_ms; the megakernel's total_ms), which puts
 * the fewest markers on the stream. Resets the statistics. */
pt_status pt_tracer_set_profiling(pt_tracer* tracer, int level);
/* Synchronises and returns the statistics accumulated since set_profiling. */
pt_status pt_tracer_stats(pt_tracer* tracer, pt_trace_stats* out);

/* ---------------------------------------------------------------- batching */
/* Wavefront tuning of a context, used by every wavefront render on it. */
typedef struct pt_wave_config {
    uint32_t lds_stack_depth;  /* traversal stack entries kept in LDS: 16, 24 or 32 */
    uint32_t chunk_paths;      /* paths in flight per chunk (the queue tile size), 4096..2^27;
                                  0: automatic (default), 2^24, or 2^26 for a launch of at least
                                  2^28 paths while 96 GiB of device memory are free */
    uint32_t sort_key;         /* ray-queue order: 0 none, 1 direction octant, 2 hit material */
    uint32_t traversal;        /* 0: each thread traces its rays one batch of the wave at a time;
                                  1: lane refill (a lane that finishes starts its next ray at once);
                                  2: threshold refill (idle lanes start their next rays together
                                     once fewer than half of the wave is still traversing) */
    uint32_t streams;          /* 1, or 2: chunks alternate between the context's two HIP streams,
                                  each with its own path state, so one chunk's shading overlaps the other's traversal;
                                  0: automatic (2 when a render spans three chunks or more, or is at
                                  least lane_split_paths and fits two chunks) */
    uint32_t trace_waves_per_cu;  /* traversal grid: one-wave blocks per CU, 16..2048 (default 512) */
    uint32_t shade_blocks_per_cu; /* shade grid: 256-thread blocks per CU, 1..256 (default 32) */
    uint32_t shade_window;        /* queue rounds of 256 entries that a shade block splits into
                                     surface hits and misses before shading: 1, 2 or 4; 0: automatic,
                                     4 for a launch over several jobs (batch, server round), 2 for one
                                     (default 0) */
    uint32_t lane_split_paths;    /* automatic lanes: a render of at least this many paths that fits
                                     two chunks is cut into two chunks on two lanes (default 2^21) */
} pt_wave_config;
#define PT_TRAVERSAL_BATCH 0
#define PT_TRAVERSAL_REFILL 1
#define PT_TRAVERSAL_FETCH 2
#define PT_SORT_NONE 0
#define PT_SORT_OCTANT 1
#define PT_SORT_MATERIAL 2
pt_status pt_context_set_wave_config(pt_context* ctx, const pt_wave_config* config);
pt_status pt_context_get_wave_config(const pt_context* ctx, pt_wave_config* config);

/* Trace samples[i] more samples on each tracers[i] in one wavefront launch
 * sequence whose queues mix every job: scenes, tile sizes and sample counts
 * may all differ (all tracers on one context). Per tracer the result equals
 * pt_tracer_trace_all(tracers[i], samples[i]) bit for bit. Asynchronous. */
pt_status pt_trace_batch(pt_tracer* const* tracers, uint32_t count, const uint32_t* samples);

/* Autotuner over the batch: times it under wave configurations (coordinate
 * descent over the traversal mode, chunk size {2M, 4M, 8M, 16M} paths, sort
 * key {none, octant, material}, LDS stack depth {16, 24, 32}, lanes {auto, 1,
 * 2}, traversal grid {256, 512, 1024} waves per CU, shade grid {16, 32, 64}
 * blocks per CU and shade window {1, 2, 4}), keeps the fastest in the context
 * and reports it. A candidate that runs out of device memory is skipped; on
 * any other error the context keeps the best configuration found so far.
 * Every trial restarts the tracers at sample 1; they are left reset
 * (sample_count 0). */
typedef struct pt_tune_result {
    pt_wave_config best;
    float best_ms;      /* batch time under the best configuration */
    float initial_ms;   /* batch time under the configuration it started from */
    uint32_t trials;
} pt_tune_result;
pt_status pt_autotune_batch(pt_tracer* const* tracers, uint32_t count, const uint32_t* samples, pt_tune_result* out);

/* ------------------------------------------------------------------ serving */
/* A render server on a context: a worker thread that keeps scenes resident,
 * admits queued jobs while slots and the HBM budget allow, and traces them
 * with continuous batching. Each round is one wavefront chunk made of pixel
 * blocks from every active job; jobs submitted meanwhile join the next
 * round. Rounds are pipelined over two lanes (the context's two HIP streams,
 * each lane with its own path state): while one round runs, the next is built and launched on the
 * other lane, so a round's last, small bounces overlap the next round's
 * first, large ones. A job's pixels always take all of their samples within
 * one round, in sample order, so its image is bit-identical however it is
 * mixed. Job buffers are sub-allocated from an arena reserved when the
 * server is created: admission, completion and collection allocate and free
 * no device memory.
 *
 * A scene stays in use from the submission of a job over it until that job
 * is done, failed or cancelled: pt_scene_destroy and the pt_scene_update_*
 * calls return PT_ERR_STATE for it meanwhile. */
typedef struct pt_server pt_server;
typedef struct pt_server_config {
    uint64_t hbm_budget_bytes;  /* admission budget: resident scenes + admitted jobs' buffers, + the
                                   worker's path state while another job runs (a job that fits alone
                                   always runs) */
    uint32_t max_active_jobs;   /* job slots */
    uint32_t round_paths;       /* paths per round (0: the context's chunk size; 2^24 when that is automatic) */
    uint64_t job_arena_bytes;   /* device arena for job buffers, reserved at creation
                                   (0: min(hbm_budget_bytes, 4 GiB)); a request whose buffer
                                   exceeds it is rejected */
} pt_server_config;
typedef struct pt_render_request {
    pt_scene* scene;            /* a scene of the server's context */
    uint32_t width, height;     /* the whole image is one tile */
    uint32_t samples;           /* samples per pixel, from sample 1 */
    uint32_t max_bounces;       /* 0..16 */
} pt_render_request;
#define PT_JOB_QUEUED 0
#define PT_JOB_RUNNING 1
#define PT_JOB_DONE 2
#define PT_JOB_FAILED 3
#define PT_JOB_CANCELLED 4
typedef struct pt_job_info {
    int32_t state;
    uint32_t rounds;            /* rounds the job took part in */
    double submit_s;            /* steady-clock seconds */
    double admit_s;
    double done_s;              /* result ready (0 until done) */
    uint64_t footprint_bytes;   /* device bytes held by the job */
    uint64_t result_bytes;      /* bytes pt_server_result writes: width * height * 16 */
} pt_job_info;
typedef struct pt_server_stats {
    uint64_t rounds;
    uint64_t paths;             /* path samples traced */
    uint64_t jobs_done;
    uint64_t jobs_rejected;
    double busy_s;              /* time with at least one round on the device */
    uint64_t bytes_in_use;      /* admitted footprint now */
    uint32_t max_jobs_per_round;    /* most jobs that shared one round */
    uint32_t max_rounds_in_flight;  /* most rounds on the device at once (2 when lanes overlapped) */
    uint64_t arena_bytes;       /* size of the job arena */
} pt_server_stats;
pt_status pt_server_create(pt_context* ctx, const pt_server_config* config, pt_server** out);
/* Cancels queued jobs, finishes the rounds in flight, joins the worker. */
pt_status pt_server_destroy(pt_server* server);
/* Queue a job; PT_ERR_MEMORY if it can never fit the budget or the arena. Thread-safe. */
pt_status pt_server_submit(pt_server* server, const pt_render_request* request, uint64_t* job_id);
/* Wait until the job is done, failed or cancelled; PT_ERR_TIMEOUT after
 * timeout_ms; PT_ERR_INVALID if the job is unknown or was collected
 * meanwhile by another thread. */
pt_status pt_server_wait(pt_server* server, uint64_t job_id, uint32_t timeout_ms);
pt_status pt_server_info(pt_server* server, uint64_t job_id, pt_job_info* info);
/* Copy a done job's accumulation (width*height float4, the sum of its
 * samples) to host memory and release the job. `capacity` is the size of
 * `rgba` in bytes; PT_ERR_INVALID (job kept) when it is smaller than the
 * job's result_bytes. For a failed or cancelled job, return PT_ERR_DEVICE or
 * PT_ERR_CANCELLED and release it (rgba is not written); PT_ERR_STATE while
 * it is queued or running. */
pt_status pt_server_result(pt_server* server, uint64_t job_id, float* rgba, size_t capacity);
/* Cancel a job. A queued job is cancelled at once. A running job is taken
 * out of scheduling and cancelled, with its buffer released, once its rounds
 * in flight end; if they complete it, it is done instead. PT_ERR_STATE for a
 * finished job. */
pt_status pt_server_cancel(pt_server* server, uint64_t job_id);
pt_status pt_server_stats_get(pt_server* server, pt_server_stats* out);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* PT_API_H */
