// render_main.cpp - pt-render: the reference's `render` command on libptcore.
//
// Keeps the reference's flag set and semantics (src/render/headless.rs:14-127):
//   --scene --skybox-luminance --output --width --height --samples
//   --max-bounces --tile-size --camera-position X,Y,Z --camera-yaw
//   --camera-pitch --camera-fov
// and its tile loop (headless.rs:207-285): the image is traced tile by tile,
// each tile to all of its samples, at most 16 samples per trace call, then
// tone-mapped (display.wgsl) and saved as RGB8 in the format the --output
// extension names (.png, .jpg/.jpeg, .webp). --scene takes a glTF file
// (.glb, or .gltf with embedded data) with --skybox <.hdr|.exr|.png>, loaded by the
// asset front end (pt_assets.h), or one of the procedural generators
// (cornell | field | large | sized:<triangles>), or a .ptscene file (the
// built arrays, camera and sky baked in, so the camera and sky flags do not
// apply). Extra flags: --seed, --device, --schedule mega|wavefront, --pfm
// <file> (linear storage image dump), and --export-glb <file.glb>
// [--export-sky <file.exr|.hdr>] / --export-ptscene <file>, which write the
// world (pt_world_export_glb, pt_scene_save), print the render flags that
// give the same camera and sky luminance, and exit without rendering.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pt_assets.h"

#include "../../include/pt_api.h"
#include "../../include/pt_scenes.h"

namespace {

[[noreturn]] void die(const char* fmt, const char* a = "") {
    std::fprintf(stderr, "error: ");
    std::fprintf(stderr, fmt, a);
    std::fprintf(stderr, "\n");
    std::exit(1);
}
void check(pt_status s, const char* what) {
    if (s != PT_OK) {
        std::fprintf(stderr, "error: %s: %s\n", what, pt_last_error());
        std::exit(1);
    }
}
float parse_finite(const char* v, const char* flag) {
    char* end = nullptr;
    float f = std::strtof(v, &end);
    if (!end || *end || !std::isfinite(f)) die("%s must be a finite number", flag);
    return f;
}
uint32_t parse_u32(const char* v, const char* flag, uint32_t lo, uint32_t hi) {
    char* end = nullptr;
    unsigned long x = std::strtoul(v, &end, 10);
    if (!end || *end || x < lo || x > hi) die("%s out of range", flag);
    return (uint32_t)x;
}

void write_pfm(const std::string& path, const std::vector<float>& rgb, uint32_t w, uint32_t h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) die("cannot write %s", path.c_str());
    std::fprintf(f, "PF\n%u %u\n-1.0\n", w, h);
    for (int64_t y = (int64_t)h - 1; y >= 0; --y) std::fwrite(&rgb[(size_t)y * w * 3], 4, (size_t)w * 3, f);
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    std::string scene = "cornell", skybox, output = "render.png", pfm, schedule = "wavefront";
    std::string export_glb, export_sky, export_ptscene;
    uint32_t width = 1920, height = 1080, samples = 256, max_bounces = PT_DEFAULT_MAX_BOUNCES, tile = 512, device = 0;
    uint64_t seed = 1;
    float sky_lum = -1.0f;
    float cam_pos[3] = {NAN, NAN, NAN};
    float yaw = NAN, pitch = NAN, fov = NAN;
    int first = 1;
    if (argc > 1 && std::strcmp(argv[1], "render") == 0) first = 2;
    for (int i = first; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) die("%s needs a value", a.c_str());
            return argv[++i];
        };
        if (a == "--scene") scene = val();
        else if (a == "--skybox") skybox = val();
        else if (a == "--output") {
            output = val();
            pt_image_format f;  // rejected here, before any work, as parse_output does (headless.rs:118-127)
            check(pt_image_format_from_path(output.c_str(), &f), "--output");
        }
        else if (a == "--pfm") pfm = val();
        else if (a == "--export-glb") export_glb = val();
        else if (a == "--export-sky") export_sky = val();
        else if (a == "--export-ptscene") export_ptscene = val();
        else if (a == "--width") width = parse_u32(val(), "--width", 1, 1u << 16);
        else if (a == "--height") height = parse_u32(val(), "--height", 1, 1u << 16);
        else if (a == "--samples") samples = parse_u32(val(), "--samples", 1, 1u << 24);
        else if (a == "--max-bounces") max_bounces = parse_u32(val(), "--max-bounces", 0, PT_MAX_BOUNCES_LIMIT);
        else if (a == "--tile-size") tile = parse_u32(val(), "--tile-size", 1, 1u << 16);
        else if (a == "--seed") seed = std::strtoull(val(), nullptr, 10);
        else if (a == "--device") device = parse_u32(val(), "--device", 0, 64);
        else if (a == "--schedule") schedule = val();
        else if (a == "--skybox-luminance") {
            sky_lum = parse_finite(val(), "--skybox-luminance");
            if (sky_lum < 0.0f) die("--skybox-luminance must be nonnegative");
        } else if (a == "--camera-position") {
            std::string v = val();
            float xyz[3];
            size_t p = 0;
            for (int k = 0; k < 3; ++k) {
                size_t q = v.find(',', p);
                if ((k < 2) != (q != std::string::npos)) die("--camera-position must be three comma-separated coordinates");
                xyz[k] = parse_finite(v.substr(p, q == std::string::npos ? std::string::npos : q - p).c_str(), "--camera-position");
                p = q + 1;
            }
            for (int k = 0; k < 3; ++k) cam_pos[k] = xyz[k];
        } else if (a == "--camera-yaw") yaw = parse_finite(val(), "--camera-yaw");
        else if (a == "--camera-pitch") {
            pitch = parse_finite(val(), "--camera-pitch");
            if (pitch < -90.0f || pitch > 90.0f) die("--camera-pitch must be from -90 to 90");
        } else if (a == "--camera-fov") {
            fov = parse_finite(val(), "--camera-fov");
            if (!(fov > 0.0f && fov < 180.0f)) die("--camera-fov must be strictly between 0 and 180");
        } else if (a == "--help" || a == "-h") {
            std::printf("usage: pt-render [render] --scene file.glb|cornell|field|large|sized:N [--skybox sky.hdr]\n"
                        "  [--width W] [--height H]\n"
                        "  [--samples S] [--max-bounces D] [--tile-size T] [--output out.png|.jpg|.webp] [--pfm out.pfm]\n"
                        "  [--skybox-luminance L] [--camera-position X,Y,Z] [--camera-yaw DEG]\n"
                        "  [--camera-pitch DEG] [--camera-fov DEG] [--seed N] [--schedule mega|wavefront]\n"
                        "  [--export-glb W.glb [--export-sky S.exr|S.hdr]] [--export-ptscene F.ptscene]\n");
            return 0;
        } else die("unknown argument %s", a.c_str());
    }
    const auto started = std::chrono::steady_clock::now();
    auto ends_with = [&](const char* suffix) {
        const size_t n = std::strlen(suffix);
        return scene.size() >= n && scene.compare(scene.size() - n, n, suffix) == 0;
    };
    if (!export_sky.empty() && export_glb.empty()) die("--export-sky goes with --export-glb");
    const bool exporting = !export_glb.empty() || !export_ptscene.empty();
    pt_generated* gen = nullptr;
    pt_assets* assets = nullptr;
    pt_scene_file* baked = nullptr;
    pt_world* world = nullptr;
    pt_world_desc wd;
    pt_scene_desc sd;
    if (ends_with(".ptscene")) {
        if (!skybox.empty() || sky_lum >= 0.0f || !std::isnan(cam_pos[0]) || !std::isnan(yaw) || !std::isnan(pitch) ||
            !std::isnan(fov))
            die("a .ptscene holds its camera and sky: --skybox, --skybox-luminance and the camera flags do not apply");
        if (!export_glb.empty()) die("--export-glb takes a world (a glTF file or a generator), not a .ptscene");
        check(pt_scene_file_load(scene.c_str(), &baked), "load .ptscene");
        check(pt_scene_file_desc(baked, &sd), ".ptscene arrays");
    } else if (ends_with(".glb") || ends_with(".gltf")) {
        if (skybox.empty()) die("--skybox <file> is required with a glTF scene");
        check(pt_assets_create(&assets), "assets");
        check(pt_assets_load_scene_file(assets, scene.c_str()), "load scene");
        check(pt_assets_load_skybox_file(assets, skybox.c_str()), "load skybox");
        pt_assets_summary info;
        check(pt_assets_info(assets, &info), "assets");
        for (uint32_t k = 0; k < info.warning_count; ++k) std::fprintf(stderr, "warning: %s\n", pt_assets_warning(assets, k));
        check(pt_assets_world_desc(assets, &wd), "scene description");
        // the headless default sky luminance (world.rs:21, headless.rs:24-31)
        wd.skybox_luminance = sky_lum >= 0.0f ? sky_lum : 8000.0f;
    } else {
        if (!skybox.empty()) die("--skybox applies to glTF scenes; the generators make their own sky");
        int kind = 0;
        uint32_t tris = 0;
        if (scene == "cornell") kind = PT_SCENE_CORNELL;
        else if (scene == "field") kind = PT_SCENE_FIELD;
        else if (scene == "large") kind = PT_SCENE_LARGE;
        else if (scene.rfind("sized:", 0) == 0) {
            kind = PT_SCENE_SIZED;
            tris = parse_u32(scene.c_str() + 6, "--scene sized:N", 64, 1u << 30);
        } else die("unknown scene %s (a .glb/.gltf file, cornell, field, large, sized:N)", scene.c_str());
        check(pt_generate_scene(kind, seed, tris, width, height, &gen), "generate scene");
        check(pt_generated_world_desc(gen, &wd), "scene description");
        if (sky_lum >= 0.0f) wd.skybox_luminance = sky_lum;
    }
    if (!baked) {
        if (!std::isnan(cam_pos[0]))
            for (int k = 0; k < 3; ++k) wd.camera_position[k] = cam_pos[k];
        if (!std::isnan(yaw)) wd.camera_yaw_degrees = yaw;
        if (!std::isnan(pitch)) wd.camera_pitch_degrees = pitch;
        if (!std::isnan(fov)) wd.camera_fov_degrees = fov;
        wd.aspect_ratio = (float)width / (float)height;
        if (!export_glb.empty()) {
            uint32_t inexact = 0;
            check(pt_world_export_glb(&wd, export_glb.c_str(), export_sky.empty() ? nullptr : export_sky.c_str(), &inexact),
                  "export glTF");
            std::printf("Wrote %s%s%s (%u normal/tangent directions inexact)\n", export_glb.c_str(),
                        export_sky.empty() ? "" : " and ", export_sky.c_str(), inexact);
            // the camera and sky luminance are render flags (headless.rs:14-64)
            std::printf("render flags:");
            if (!std::isnan(wd.camera_position[0]))
                std::printf(" --camera-position %.9g,%.9g,%.9g", wd.camera_position[0], wd.camera_position[1],
                            wd.camera_position[2]);
            if (!std::isnan(wd.camera_yaw_degrees)) std::printf(" --camera-yaw %.9g", wd.camera_yaw_degrees);
            if (!std::isnan(wd.camera_pitch_degrees)) std::printf(" --camera-pitch %.9g", wd.camera_pitch_degrees);
            if (!std::isnan(wd.camera_fov_degrees)) std::printf(" --camera-fov %.9g", wd.camera_fov_degrees);
            if (wd.skybox_luminance >= 0.0f) std::printf(" --skybox-luminance %.9g", wd.skybox_luminance);
            std::printf(" --width %u --height %u\n", width, height);
        }
        check(pt_world_build(&wd, &world), "build world");
        check(pt_world_scene_desc(world, &sd), "scene arrays");
    }
    if (!export_ptscene.empty()) {
        check(pt_scene_save(&sd, export_ptscene.c_str()), "save .ptscene");
        std::printf("Wrote %s (%u triangles, %u BVH nodes)\n", export_ptscene.c_str(), sd.triangle_count, sd.bvh_node_count);
    }
    if (exporting) {
        pt_world_destroy(world);
        pt_generated_destroy(gen);
        pt_assets_destroy(assets);
        pt_scene_file_destroy(baked);
        return 0;
    }
    std::printf("Loaded scene %s with %u triangles, %u materials, %u BVH nodes\n", scene.c_str(), sd.triangle_count,
                sd.material_count, sd.bvh_node_count);

    pt_context* ctx = nullptr;
    check(pt_context_create((int)device, &ctx), "create context");
    pt_scene* dscene = nullptr;
    check(pt_scene_create(ctx, &sd, &dscene), "upload scene");
    pt_world_destroy(world);  // the GPU has its own copy (headless.rs:157)
    pt_generated_destroy(gen);
    pt_assets_destroy(assets);
    pt_scene_file_destroy(baked);

    const uint32_t tw = std::min(tile, width), th = std::min(tile, height);
    pt_tracer* tracer = nullptr;
    check(pt_tracer_create(ctx, dscene, tw, th, &tracer), "create tracer");
    check(pt_tracer_set_schedule(tracer, schedule == "mega" ? PT_SCHEDULE_MEGAKERNEL : PT_SCHEDULE_WAVEFRONT), "schedule");
    check(pt_tracer_set_max_bounces(tracer, max_bounces), "max bounces");
    std::vector<uint8_t> rgb((size_t)width * height * 3);
    std::vector<float> lin(pfm.empty() ? 0 : (size_t)width * height * 3);
    std::vector<uint8_t> tile_rgba((size_t)tw * th * 4);
    std::vector<float> tile_f(pfm.empty() ? 0 : (size_t)tw * th * 4);
    const auto trace_start = std::chrono::steady_clock::now();
    for (uint32_t y = 0; y < height; y += th) {
        for (uint32_t x = 0; x < width; x += tw) {
            const uint32_t vw = std::min(tw, width - x), vh = std::min(th, height - y);
            check(pt_tracer_set_tile(tracer, width, height, x, y, tw, th), "set tile");
            uint32_t remaining = samples;
            while (remaining > 0) {
                uint32_t n = 0;
                check(pt_tracer_trace(tracer, remaining, &n), "trace");
                remaining -= n;
            }
            check(pt_tracer_display_rgba8(tracer, tile_rgba.data()), "display");
            if (!pfm.empty()) check(pt_tracer_read_storage(tracer, tile_f.data()), "read storage");
            for (uint32_t r = 0; r < vh; ++r)
                for (uint32_t c = 0; c < vw; ++c)
                    for (int k = 0; k < 3; ++k) {
                        rgb[((size_t)(y + r) * width + x + c) * 3 + k] = tile_rgba[((size_t)r * tw + c) * 4 + k];
                        if (!pfm.empty()) lin[((size_t)(y + r) * width + x + c) * 3 + k] = tile_f[((size_t)r * tw + c) * 4 + k];
                    }
        }
    }
    const double trace_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - trace_start).count();
    check(pt_image_save(output.c_str(), rgb.data(), width, height), "save");
    if (!pfm.empty()) write_pfm(pfm, lin, width, height);
    const double total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    std::printf("Saved %s in %.1fs (trace %.3fs, %.3g path samples/s)\n", output.c_str(), total_s, trace_s,
                (double)width * height * samples / trace_s);
    pt_tracer_destroy(tracer);
    pt_scene_destroy(dscene);
    pt_context_destroy(ctx);
    return 0;
}
