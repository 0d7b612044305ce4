/*
 * pt_assets.h - scene and sky files to the input of pt_world_build.
 *
 * Replaces the loading half of World::new (src/render/world.rs:589-611):
 * Scene::new for glTF 2.0 scenes (world.rs:887-1423: images to RGBA8 texture
 * atlases, node flattening with transforms, normals, MikkTSpace tangents,
 * materials with OpenPBR extras) and the decode step of Skybox::new
 * (world.rs:1872-1873). The arrays come out in the reference's input order,
 * so pt_world_build then builds the BVH, alias tables and camera as the
 * reference does. Supported inputs: .glb (or .gltf with embedded data URIs)
 * with PNG or JPEG images; sky images in Radiance .hdr, scanline OpenEXR (NONE, ZIPS, ZIP, PIZ), PNG or JPEG.
 * Also the output-image writers of the `render` command (PNG, JPEG, WebP).
 */
#ifndef PT_ASSETS_H
#define PT_ASSETS_H

#include <stddef.h>
#include <stdint.h>

#include "pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_assets pt_assets;

typedef struct pt_assets_summary {
    uint32_t triangle_count;
    uint32_t material_count;
    uint32_t srgb_atlas_count;   /* 8192^2 RGBA8 layers */
    uint32_t linear_atlas_count;
    uint32_t atlas_size;         /* 8192 when there are layers, else 0 */
    uint32_t skybox_width;
    uint32_t skybox_height;
    uint32_t warning_count;      /* messages the reference would log (skipped primitives, extras) */
} pt_assets_summary;

pt_status pt_assets_create(pt_assets** out);
pt_status pt_assets_destroy(pt_assets* assets);
/* Scene::new on the bytes of a .glb / embedded .gltf; replaces a previous scene. */
pt_status pt_assets_load_scene(pt_assets* assets, const uint8_t* bytes, size_t size);
pt_status pt_assets_load_scene_file(pt_assets* assets, const char* path);
/* image::load_from_memory(..).into_rgba32f() of a sky image; replaces a previous sky. */
pt_status pt_assets_load_skybox(pt_assets* assets, const uint8_t* bytes, size_t size);
pt_status pt_assets_load_skybox_file(pt_assets* assets, const char* path);
pt_status pt_assets_info(const pt_assets* assets, pt_assets_summary* out);
/* Warning `index` (< warning_count), or "" . */
const char* pt_assets_warning(const pt_assets* assets, uint32_t index);
/* A pt_world_desc borrowing the loaded arrays (valid until the next load or
 * destroy): camera fields NaN (autogenerated), skybox_luminance -1 (the
 * image's own), aspect ratio 1; the caller sets the render options. */
pt_status pt_assets_world_desc(const pt_assets* assets, pt_world_desc* out);

/* Output images: the saving step of `render` (src/render/headless.rs:118-127,
 * 163-170), image::save_buffer(path, rgb8, width, height, Rgb8) with the
 * format taken from the file extension. PNG and WebP are lossless; JPEG is
 * baseline with 4:4:4 sampling at quality 75 unless asked otherwise. */
typedef enum pt_image_format {
    PT_IMAGE_PNG = 0,
    PT_IMAGE_JPEG = 1,
    PT_IMAGE_WEBP = 2, /* lossless (VP8L) */
} pt_image_format;

/* image::ImageFormat::from_path restricted to the formats `render` accepts
 * (parse_output, headless.rs:118-127): .png, .jpg, .jpeg, .webp, any case. */
pt_status pt_image_format_from_path(const char* path, pt_image_format* format);
/* Encodes width x height RGB8 pixels (rows row_stride bytes apart, 0 = width * 3).
 * quality: JPEG 1..100, 0 = 75; ignored by the lossless formats. *out is
 * allocated by the library and released with pt_image_free. */
pt_status pt_image_encode(pt_image_format format, const uint8_t* rgb, size_t row_stride, uint32_t width,
                          uint32_t height, uint32_t quality, uint8_t** out, size_t* size);
void pt_image_free(uint8_t* data);
/* image::save_buffer: encode by extension and write the file. */
pt_status pt_image_save(const char* path, const uint8_t* rgb, uint32_t width, uint32_t height);

/* ------------------------------------------------------------- scene export */
/* Write a world's input (what pt_world_build takes) as a glTF binary that
 * the reference's Scene::new (src/render/world.rs:887-1423) and
 * pt_assets_load_scene read back into the same triangles, materials and
 * atlases, and (sky_path may be NULL) its sky as an uncompressed FLOAT
 * OpenEXR that Skybox::new (world.rs:1872-1873) and pt_assets_load_skybox
 * read back exactly; a sky_path ending in ".hdr" writes Radiance RGBE
 * instead, and fails (PT_ERR_INVALID) unless every sky value is exactly an
 * RGBE value (c * 2^(e - 136), alpha 1).
 * One node without a transform holds one mesh with a primitive per run of
 * triangles of one material, with explicit NORMAL, TANGENT and TEXCOORD_0
 * per corner; textures become RGBA PNG images. Each normal and tangent
 * direction is written as a vector that the loader's normalisation turns
 * back into the stored one; where none lies within 3 ulps it is written as
 * stored and loads back within a few ulps, and *inexact_directions (may be
 * NULL) counts those (0 for worlds from pt_generate_scene and
 * pt_assets_load_scene). Materials not numbered in order of first use are
 * written but load renumbered. PT_ERR_INVALID for what the format cannot
 * hold: atlas layers other than 8192^2, texture rectangles the loader's
 * packer would place elsewhere, OpenPBR parameters outside [0, 1] and a
 * normal scale without a normal texture. The camera and the sky luminance
 * are render flags, not file contents. */
pt_status pt_world_export_glb(const pt_world_desc* desc, const char* glb_path, const char* sky_path,
                              uint32_t* inexact_directions);

/* The device-layout arrays of a scene (pt_scene_desc, BVH included) as a
 * .ptscene file, and back. No counterpart in the reference: it lets a tool
 * hand over the identical serialized BVH. A loaded file owns its arrays;
 * pt_scene_file_desc borrows them until pt_scene_file_destroy. Loading
 * refuses (PT_ERR_INVALID) a damaged file and one whose arrays
 * pt_scene_create would refuse (an index the kernels follow out of range, a
 * BVH that is not a tree), so it can be checked without a device. */
typedef struct pt_scene_file pt_scene_file;
pt_status pt_scene_save(const pt_scene_desc* desc, const char* path);
pt_status pt_scene_file_load(const char* path, pt_scene_file** out);
pt_status pt_scene_file_desc(const pt_scene_file* file, pt_scene_desc* out);
pt_status pt_scene_file_destroy(pt_scene_file* file);

#ifdef __cplusplus
}
#endif

#endif /* PT_ASSETS_H */
