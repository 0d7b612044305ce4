/*
 * pt_scenes.h - seeded procedural scenes for the benchmark workloads.
 *
 * No counterpart in the reference (it renders .glb files); these produce the
 * same kind of input World::new would get after glTF flattening: a triangle
 * soup in input order, materials, texture atlases and an equirectangular sky,
 * as a pt_world_desc for pt_world_build.
 *
 *   PT_SCENE_CORNELL   Cornell box: 5 walls (10 tris), two boxes rotated
 *                      +-17 degrees (24 tris), emissive ceiling quad (2 tris,
 *                      radiance 15), diffuse EON albedos in [0.1, 0.8],
 *                      constant grey sky. 36 triangles.
 *   PT_SCENE_FIELD     heightfield of value noise (amplitude 0.2) plus 64
 *                      icospheres of 1280 triangles; GGX roughness U[0.05, 0.9],
 *                      metalness Bernoulli(0.3); 256 emissive triangles with flux
 *                      log-uniform over [1, 1000]; gradient sky. The grid is
 *                      sized so the total is ~262k triangles (301^2 vertices:
 *                      180,000 + 81,920 = 261,920).
 *   PT_SCENE_LARGE     heightfield (1202^2 vertices, 2,884,802 tris) plus 4096
 *                      icospheres of 320 triangles (1,310,720): 4,195,522 tris;
 *                      8 Perlin 2048^2 RGBA8 textures (4 sRGB albedo, 4 linear
 *                      metallic-roughness), 4096 emissive triangles, 1024x512
 *                      RGBE-quantised procedural sky.
 *   PT_SCENE_SIZED     heightfield + icospheres with about `triangles` triangles
 *                      (the per-scene generator of the mixed batch workload).
 */
#ifndef PT_SCENES_H
#define PT_SCENES_H

#include "pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PT_SCENE_CORNELL 1
#define PT_SCENE_FIELD 2
#define PT_SCENE_LARGE 3
#define PT_SCENE_SIZED 4

typedef struct pt_generated pt_generated;

/* Generate scene `kind` from `seed`. `triangles` is used by PT_SCENE_SIZED
 * only. width/height set the camera aspect ratio. */
pt_status pt_generate_scene(int kind, uint64_t seed, uint32_t triangles, uint32_t width, uint32_t height,
                            pt_generated** out);
/* Borrow the generated arrays (valid until pt_generated_destroy). */
pt_status pt_generated_world_desc(const pt_generated* g, pt_world_desc* out);
pt_status pt_generated_destroy(pt_generated* g);

#ifdef __cplusplus
}
#endif

#endif /* PT_SCENES_H */
