// api.cpp - the C ABI of libptcore.so: contexts, device scenes, tracers.
//
// Mirrors the reference's host objects around its trace kernel:
//   pt_context  <- wgpu adapter/device/queue (gpu.rs:14-76): one HIP device + two streams
//   pt_scene    <- WorldResources (world_resources.rs:74-245, 479-525, 600-661)
//   pt_tracer   <- PathTracer (path_tracer.rs:59-358)
// Limits the reference enforces with anyhow errors (tile size, buffer sizes,
// world_resources.rs:606-642, path_tracer.rs:361-382) become PT_ERR_INVALID;
// in addition, every index the kernels follow is validated here once, because
// the kernels (like the reference's, path_tracer.rs:169-182) do no bounds checks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"
#include "pt_internal.h"
#include "pt_runtime.h"
#include "wavefront.h"

namespace ptcore {

static thread_local std::string g_error;

// Messages can quote bytes of the caller's files (a corrupt header line, a
// JSON string). Whatever is not well-formed UTF-8 (including a sequence cut
// by the length limit) becomes '?', so the message is always valid text.
std::string utf8_clean(const char* s) {
    std::string out;
    const unsigned char* p = (const unsigned char*)s;
    while (*p) {
        const unsigned char c = *p;
        int n = c < 0x80 ? 0 : (c >> 5) == 6 ? 1 : (c >> 4) == 14 ? 2 : (c >> 3) == 30 ? 3 : -1;
        bool ok = n >= 0;
        for (int k = 1; ok && k <= n; ++k) ok = (p[k] & 0xC0) == 0x80;
        if (ok && n == 1) ok = c >= 0xC2;                                  // no overlong 2-byte forms
        if (ok && n == 2) ok = !(c == 0xE0 && p[1] < 0xA0) && !(c == 0xED && p[1] >= 0xA0);  // overlong, surrogates
        if (ok && n == 3) ok = c <= 0xF4 && !(c == 0xF0 && p[1] < 0x90) && !(c == 0xF4 && p[1] >= 0x90);
        if (ok && (c >= 0x20 || c == '\n' || c == '\t')) {
            out.append((const char*)p, (size_t)n + 1);
            p += n + 1;
        } else {
            out.push_back('?');
            ++p;
        }
    }
    return out;
}

pt_status set_error(pt_status code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = utf8_clean(buf);
    return code;
}

pt_status hip_error(hipError_t e, const char* what) {
    return set_error(e == hipErrorOutOfMemory ? PT_ERR_MEMORY : PT_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

void srgb_table(float out[256]) {
    // Rgba8UnormSrgb decode (world_resources.rs:90): IEC 61966-2-1 EOTF
    for (int i = 0; i < 256; ++i) {
        double c = i / 255.0;
        out[i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
}

pt_status scene_require_idle(const Scene* s, const char* what) {
    const uint32_t n = s->server_jobs.load();
    if (n) return set_error(PT_ERR_STATE, "%s: %u render-server job(s) over this scene are not finished", what, n);
    return PT_OK;
}

pt_status Context::ensure_overflow(uint32_t depth, uint32_t grid_threads) {
    uint32_t extra = depth > ptd::MEGA_LDS_DEPTH ? depth - ptd::MEGA_LDS_DEPTH : 1u;
    size_t need = (size_t)extra * grid_threads * sizeof(uint32_t);
    if (need <= ovf_bytes) return PT_OK;
    if (ovf) {
        hipError_t se = hipStreamSynchronize(stream);
        if (se != hipSuccess) return hip_error(se, "traversal overflow stack");
        (void)hipFree(ovf);
        ovf = nullptr;
        ovf_bytes = 0;
    }
    hipError_t e = hipMalloc(&ovf, need);
    if (e != hipSuccess) return hip_error(e, "allocating the traversal overflow stack");
    ovf_bytes = need;
    return PT_OK;
}

}  // namespace ptcore

using namespace ptcore;

namespace {

template <class T>
pt_status upload(T** dst, const T* src, size_t count, size_t& footprint, const char* what) {
    *dst = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return hip_error(e, what);
    footprint += bytes;
    if (count && src) {
        e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_error(e, what);
    } else {
        e = hipMemset(*dst, 0, bytes);
        if (e != hipSuccess) return hip_error(e, what);
    }
    return PT_OK;
}

}  // namespace

namespace ptcore {

// BVH validity and depth: every child pair, leaf range and node is in range,
// and the tree is acyclic (depth bounded by the node count).
static pt_status validate_bvh(const pt_scene_desc* d, uint32_t& max_depth) {
    if (d->bvh_node_count == 0 || !d->bvh_nodes) return set_error(PT_ERR_INVALID, "scene has no BVH nodes");
    max_depth = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    stack.push_back({0u, 0u});
    uint64_t visited = 0;
    while (!stack.empty()) {
        auto [ni, depth] = stack.back();
        stack.pop_back();
        if (++visited > d->bvh_node_count) return set_error(PT_ERR_INVALID, "BVH has a cycle");
        max_depth = std::max(max_depth, depth);
        const PtBvhNode& n = d->bvh_nodes[ni];
        if (n.tri_count > 0) {
            if ((uint64_t)n.left_first + n.tri_count > d->triangle_count)
                return set_error(PT_ERR_INVALID, "BVH leaf %u indexes triangles past the end", ni);
            if (n.tri_count <= 2 && n.left_first >= (1u << 30))
                return set_error(PT_ERR_INVALID, "BVH leaf %u: triangle index too large", ni);
        } else {
            if (n.left_first == 0 || (uint64_t)n.left_first + 1 >= d->bvh_node_count || n.left_first >= (1u << 30))
                return set_error(PT_ERR_INVALID, "BVH node %u has children out of range", ni);
            stack.push_back({n.left_first + 1, depth + 1});
            stack.push_back({n.left_first, depth + 1});
        }
    }
    if (d->bvh_node_count >= (1u << 30)) return set_error(PT_ERR_INVALID, "too many BVH nodes");
    return PT_OK;
}

pt_status validate_scene(const pt_scene_desc* d, uint32_t& depth) {
    if (d->triangle_count == 0 || !d->geometries || !d->attributes)
        return set_error(PT_ERR_INVALID, "scene has no triangles");
    if (d->material_count == 0 || !d->materials) return set_error(PT_ERR_INVALID, "scene has no materials");
    for (uint32_t i = 0; i < d->triangle_count; ++i)
        if (d->attributes[i].material >= d->material_count)
            return set_error(PT_ERR_INVALID, "triangle %u uses material %u of %u", i, d->attributes[i].material,
                             d->material_count);
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const PtMaterial& m = d->materials[i];
        const int32_t s[3] = {m.base_color_tex_layer, m.emission_tex_layer, m.specular_color_tex_layer};
        const int32_t l[3] = {m.normal_tex_layer, m.metallic_roughness_tex_layer, m.specular_weight_tex_layer};
        for (int k = 0; k < 3; ++k) {
            if (s[k] >= (int32_t)d->srgb_atlas_count || l[k] >= (int32_t)d->linear_atlas_count)
                return set_error(PT_ERR_INVALID, "material %u names a missing atlas layer", i);
        }
    }
    if ((d->srgb_atlas_count || d->linear_atlas_count) && d->atlas_size == 0)
        return set_error(PT_ERR_INVALID, "atlas_size is 0");
    if (d->srgb_atlas_count && !d->srgb_atlases) return set_error(PT_ERR_INVALID, "missing sRGB atlases");
    if (d->linear_atlas_count && !d->linear_atlases) return set_error(PT_ERR_INVALID, "missing linear atlases");
    if (d->skybox.width == 0 || d->skybox.height == 0 || !d->skybox_rgba)
        return set_error(PT_ERR_INVALID, "missing skybox");
    const uint64_t texels = (uint64_t)d->skybox.width * d->skybox.height;
    if (texels > 0xffffffffull) return set_error(PT_ERR_INVALID, "skybox too large");
    const float pg = d->scene_uniforms.light_probabilities[0], ps = d->scene_uniforms.light_probabilities[1];
    // the sky branch of sample_light is reachable whenever the geometry
    // category is not certain (trace.slang:539, 564)
    if (ps > 0.0f || (pg + ps > 0.0f && pg < 1.0f)) {
        if (d->skybox_alias_count != texels || !d->skybox_alias)
            return set_error(PT_ERR_INVALID, "sky sampling needs one alias entry per sky texel");
        for (uint32_t i = 0; i < d->skybox_alias_count; ++i)
            if (d->skybox_alias[i].alias_index >= texels) return set_error(PT_ERR_INVALID, "sky alias %u out of range", i);
    }
    const uint32_t ne = d->scene_uniforms.emissive_triangles_count;
    if (ne > d->scene_alias_count || (ne && !d->scene_alias))
        return set_error(PT_ERR_INVALID, "scene alias table shorter than emissive_triangles_count");
    if (pg > 0.0f && ne == 0) return set_error(PT_ERR_INVALID, "geometry light probability without emitters");
    for (uint32_t i = 0; i < ne; ++i) {
        const PtSceneAliasEntry& e = d->scene_alias[i];
        if (e.alias_index >= ne || e.triangle_index >= d->triangle_count)
            return set_error(PT_ERR_INVALID, "scene alias entry %u out of range", i);
        if (i && d->scene_alias[i - 1].triangle_index >= e.triangle_index)
            return set_error(PT_ERR_INVALID, "scene alias entries must be sorted by triangle index");
    }
    return validate_bvh(d, depth);
}

}  // namespace ptcore

extern "C" {

const char* pt_last_error(void) { return g_error.c_str(); }
const char* pt_version(void) { return "ptcore 0.1 (gfx950)"; }

pt_status pt_context_create(int device, pt_context** out) {
    if (!out) return set_error(PT_ERR_INVALID, "pt_context_create: null output");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return set_error(PT_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return set_error(PT_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_error(e, "hipSetDevice");
    pt_context* c = new pt_context();
    c->device = device;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return hip_error(e, "hipGetDeviceProperties");
    }
    c->cu_count = (uint32_t)prop.multiProcessorCount;
    c->clock_khz = (uint32_t)prop.clockRate;
    c->l2_bytes = (uint32_t)prop.l2CacheSize;
    c->hbm_bytes = (uint64_t)prop.totalGlobalMem;
    std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->lane_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return hip_error(e, "hipStreamCreate");
    }
    float lut[256];
    srgb_table(lut);
    e = hipMalloc(&c->srgb_lut, sizeof(lut));
    if (e == hipSuccess) e = hipMemcpy(c->srgb_lut, lut, sizeof(lut), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&c->counters, 16 * sizeof(unsigned long long));
    if (e == hipSuccess) e = noalloc_scratch_address(device, &c->noalloc_scratch);
    if (e != hipSuccess) {
        pt_context_destroy(c);
        return hip_error(e, "context buffers");
    }
    *out = c;
    return PT_OK;
}

// Destruction releases everything; a failure of the work still pending on
// the stream is reported after the release (the releases' own results are not).
pt_status pt_context_destroy(pt_context* c) {
    if (!c) return PT_OK;
    (void)hipSetDevice(c->device);
    hipError_t e = c->stream ? hipStreamSynchronize(c->stream) : hipSuccess;
    const hipError_t el = c->lane_stream ? hipStreamSynchronize(c->lane_stream) : hipSuccess;
    if (e == hipSuccess) e = el;
    (void)hipFree(c->ovf);
    (void)hipFree(c->srgb_lut);
    (void)hipFree(c->counters);
    delete c->batch_wave;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->lane_stream) (void)hipStreamDestroy(c->lane_stream);
    delete c;
    return e == hipSuccess ? PT_OK : hip_error(e, "work pending at context destruction");
}

pt_status pt_context_get_info(const pt_context* c, pt_context_info* info) {
    if (!c || !info) return set_error(PT_ERR_INVALID, "pt_context_get_info: null argument");
    std::memset(info, 0, sizeof(*info));
    info->device = c->device;
    info->cu_count = c->cu_count;
    info->clock_khz = c->clock_khz;
    info->l2_bytes = c->l2_bytes;
    info->hbm_bytes = c->hbm_bytes;
    std::snprintf(info->arch, sizeof(info->arch), "%s", c->arch);
    return PT_OK;
}

pt_status pt_context_measure_gather(pt_context* c, uint32_t lanes_per_segment, double* rate) {
    if (!c || !rate) return set_error(PT_ERR_INVALID, "pt_context_measure_gather: null argument");
    if (lanes_per_segment != 1 && lanes_per_segment != 4)
        return set_error(PT_ERR_INVALID, "lanes_per_segment must be 1 or 4");
    if (pt_status st = use_device(c->device)) return st;
    const uint32_t blocks = c->cu_count * 64u;
    DeviceBuffer data, out;
    if (pt_status st = data.alloc(gather_probe_data_bytes())) return st;
    if (pt_status st = out.alloc((size_t)blocks * 64u * sizeof(float))) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemsetAsync(data.ptr, 0, data.bytes, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    constexpr int TIMED = 5;
    const uint32_t shared = lanes_per_segment == 4 ? 1u : 0u;
    if (e == hipSuccess) e = launch_gather_probe((const float4*)data.ptr, shared, (float*)out.ptr, blocks, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int i = 0; i < TIMED && e == hipSuccess; ++i)
        e = launch_gather_probe((const float4*)data.ptr, shared, (float*)out.ptr, blocks, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_error(e, "gather probe");
    const double cycles = (double)ms * 1e-3 * (double)c->clock_khz * 1e3;
    if (cycles <= 0.0 || c->cu_count == 0) return set_error(PT_ERR_DEVICE, "gather probe: no elapsed time");
    *rate = (double)gather_probe_lane_accesses(blocks) * TIMED / ((double)c->cu_count * cycles);
    return PT_OK;
}

pt_status pt_context_check_fast_math(pt_context* c, uint32_t sqrt_first, uint32_t sqrt_last, uint64_t div_pairs,
                                     uint32_t div_seed, uint64_t* sqrt_mismatches, uint64_t* div_mismatches) {
    if (!c || !sqrt_mismatches || !div_mismatches)
        return set_error(PT_ERR_INVALID, "pt_context_check_fast_math: null argument");
    if (sqrt_last < sqrt_first) return set_error(PT_ERR_INVALID, "pt_context_check_fast_math: sqrt_last < sqrt_first");
    if (pt_status st = use_device(c->device)) return st;
    DeviceBuffer counts;
    if (pt_status st = counts.alloc(2 * sizeof(unsigned long long))) return st;
    unsigned long long* d = (unsigned long long*)counts.ptr;
    const uint32_t blocks = std::max(1u, c->cu_count) * 16u;
    hipError_t e = hipMemsetAsync(d, 0, counts.bytes, c->stream);
    if (e == hipSuccess)
        e = launch_check_sqrt(sqrt_first, (uint64_t)sqrt_last - sqrt_first + 1u, d, blocks, c->stream);
    if (e == hipSuccess && div_pairs) e = launch_check_div(div_pairs, div_seed, d + 1, blocks, c->stream);
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_error(e, "fast math check");
    *sqrt_mismatches = h[0];
    *div_mismatches = h[1];
    return PT_OK;
}

pt_status pt_context_synchronize(pt_context* c) {
    if (!c) return set_error(PT_ERR_INVALID, "null context");
    hipError_t e = hipStreamSynchronize(c->stream);
    const hipError_t el = hipStreamSynchronize(c->lane_stream);
    if (e == hipSuccess) e = el;
    return e == hipSuccess ? PT_OK : hip_error(e, "hipStreamSynchronize");
}

pt_status pt_scene_create(pt_context* c, const pt_scene_desc* d, pt_scene** out) {
    if (!c || !d || !out) return set_error(PT_ERR_INVALID, "pt_scene_create: null argument");
    *out = nullptr;
    uint32_t depth = 0;
    pt_status st = validate_scene(d, depth);
    if (st != PT_OK) return st;
    if (pt_status dst = use_device(c->device)) return dst;
    pt_scene* s = new pt_scene();
    s->ctx = c;
    s->bvh_depth = depth;
    s->triangle_count = d->triangle_count;
    s->material_count = d->material_count;
    s->bvh_node_count = d->bvh_node_count;
    // like the reference, room for an alias entry per triangle so light
    // updates never reallocate (world_resources.rs:158-165)
    s->scene_alias_capacity = std::max(d->scene_alias_count, d->triangle_count);
    s->alias_count = d->scene_alias_count;
    size_t fp = 0;
    ptd::DevScene& v = s->view;
    std::memset(&v, 0, sizeof(v));
    const size_t layer = (size_t)d->atlas_size * d->atlas_size * 4;
    std::vector<uint8_t> opaque(d->triangle_count);
    bool all_opaque = true;
    for (uint32_t i = 0; i < d->triangle_count; ++i) {
        opaque[i] = d->materials[d->attributes[i].material].alpha_mode == PT_ALPHA_MODE_OPAQUE ? 1 : 0;
        all_opaque = all_opaque && opaque[i];
    }
    // traversal node copy with each node's child code precomputed (see DevScene::tnode).
    // Child pairs start at odd node indices (the root is alone at 0), so a
    // 32-byte lead puts every 64-byte pair record in one 64-byte segment of
    // the 256-byte-aligned allocation: one L1 data-return beat per record.
    std::vector<float4> tnodes(ptd::TNODE_LEAD + 2 * (size_t)d->bvh_node_count, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (uint32_t i = 0; i < d->bvh_node_count; ++i) {
        const PtBvhNode& n = d->bvh_nodes[i];
        uint32_t code = ptd::child_code(i, n.left_first, n.tri_count);
        float cf, tf;
        std::memcpy(&cf, &code, 4);
        std::memcpy(&tf, &n.tri_count, 4);
        tnodes[ptd::TNODE_LEAD + 2 * (size_t)i] = make_float4(n.aabb_min[0], n.aabb_min[1], n.aabb_min[2], cf);
        tnodes[ptd::TNODE_LEAD + 2 * (size_t)i + 1] = make_float4(n.aabb_max[0], n.aabb_max[1], n.aabb_max[2], tf);
    }
    std::vector<float4> sky((size_t)d->skybox.width * d->skybox.height);
    std::memcpy(sky.data(), d->skybox_rgba, sky.size() * sizeof(float4));
    PtTriangleGeometry* geo = nullptr;
    PtTriangleAttribute* attr = nullptr;
    PtMaterial* mat = nullptr;
    PtBvhNode* bvh = nullptr;
    float4* tn = nullptr;
    uint8_t *srgb = nullptr, *lin = nullptr, *opq = nullptr;
    float4* skyd = nullptr;
    PtSkyboxAliasEntry* sa = nullptr;
    PtSceneAliasEntry* ca = nullptr;
    auto keep_pointers = [&] {  // the view owns whatever was allocated, so destroy frees it
        v.geo = geo;
        v.attr = attr;
        v.mat = mat;
        v.bvh = bvh;
        s->tnode_alloc = tn;
        v.tnode = tn ? tn + ptd::TNODE_LEAD : nullptr;
        v.srgb = srgb;
        v.linear = lin;
        v.tri_opaque = opq;
        v.sky = skyd;
        v.sky_alias = sa;
        v.scene_alias = ca;
    };
    if ((st = upload(&geo, d->geometries, d->triangle_count, fp, "geometries")) ||
        (st = upload(&attr, d->attributes, d->triangle_count, fp, "attributes")) ||
        (st = upload(&mat, d->materials, d->material_count, fp, "materials")) ||
        (st = upload(&bvh, d->bvh_nodes, d->bvh_node_count, fp, "bvh nodes")) ||
        (st = upload(&tn, tnodes.data(), tnodes.size(), fp, "traversal nodes")) ||
        (st = upload(&srgb, d->srgb_atlases, layer * d->srgb_atlas_count, fp, "srgb atlases")) ||
        (st = upload(&lin, d->linear_atlases, layer * d->linear_atlas_count, fp, "linear atlases")) ||
        (st = upload(&opq, opaque.data(), opaque.size(), fp, "opacity")) ||
        (st = upload(&skyd, sky.data(), sky.size(), fp, "skybox")) ||
        (st = upload(&sa, d->skybox_alias, d->skybox_alias_count, fp, "skybox alias")) ||
        (st = upload(&ca, (const PtSceneAliasEntry*)nullptr, s->scene_alias_capacity, fp, "scene alias"))) {
        keep_pointers();
        pt_scene_destroy(s);
        return st;
    }
    if (d->scene_alias_count) {
        hipError_t e = hipMemcpy(ca, d->scene_alias, d->scene_alias_count * sizeof(PtSceneAliasEntry), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            keep_pointers();
            pt_scene_destroy(s);
            return hip_error(e, "scene alias");
        }
    }
    keep_pointers();
    v.srgb_lut = c->srgb_lut;
    v.camera = d->camera;
    v.skybox = d->skybox;
    v.uniforms = d->scene_uniforms;
    v.atlas_size = d->atlas_size;
    v.triangle_count = d->triangle_count;
    v.all_opaque = all_opaque ? 1u : 0u;
    s->skybox_alias_count = d->skybox_alias_count;
    s->footprint = fp;
    *out = s;
    return PT_OK;
}

pt_status pt_scene_destroy(pt_scene* s) {
    if (!s) return PT_OK;
    if (pt_status st = scene_require_idle(s, "pt_scene_destroy")) return st;
    (void)hipSetDevice(s->ctx->device);
    const hipError_t pending = hipStreamSynchronize(s->ctx->stream);
    const ptd::DevScene& v = s->view;
    (void)hipFree((void*)v.geo);
    (void)hipFree((void*)v.attr);
    (void)hipFree((void*)v.mat);
    (void)hipFree((void*)v.bvh);
    (void)hipFree((void*)s->tnode_alloc);
    (void)hipFree((void*)v.srgb);
    (void)hipFree((void*)v.linear);
    (void)hipFree((void*)v.tri_opaque);
    (void)hipFree((void*)v.sky);
    (void)hipFree((void*)v.sky_alias);
    (void)hipFree((void*)v.scene_alias);
    delete s;
    return pending == hipSuccess ? PT_OK : hip_error(pending, "work pending at scene destruction");
}

pt_status pt_scene_footprint_bytes(const pt_scene* s, uint64_t* bytes) {
    if (!s || !bytes) return set_error(PT_ERR_INVALID, "null argument");
    *bytes = s->footprint;
    return PT_OK;
}

pt_status pt_scene_update_camera(pt_scene* s, const PtCamera* camera) {
    if (!s || !camera) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status st = scene_require_idle(s, "pt_scene_update_camera")) return st;
    s->view.camera = *camera;
    ++s->version;
    return PT_OK;
}

pt_status pt_scene_update_materials(pt_scene* s, const uint32_t* idx, const PtMaterial* m, uint32_t n) {
    if (!s || (n && (!idx || !m))) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status st = scene_require_idle(s, "pt_scene_update_materials")) return st;
    for (uint32_t i = 0; i < n; ++i) {
        if (idx[i] >= s->material_count) return set_error(PT_ERR_INVALID, "material index %u out of range", idx[i]);
        if (m[i].alpha_mode != PT_ALPHA_MODE_OPAQUE && s->view.all_opaque)
            return set_error(PT_ERR_INVALID, "switching a material to alpha-tested needs a new scene");
    }
    if (pt_status dst = use_device(s->ctx->device)) return dst;
    for (uint32_t i = 0; i < n; ++i) {
        hipError_t e = hipMemcpyAsync((void*)(s->view.mat + idx[i]), &m[i], sizeof(PtMaterial), hipMemcpyHostToDevice,
                                      s->ctx->stream);
        if (e != hipSuccess) return hip_error(e, "material update");
    }
    hipError_t se = hipStreamSynchronize(s->ctx->stream);
    if (se != hipSuccess) return hip_error(se, "material update");
    ++s->version;
    return PT_OK;
}

// Checks that light uniforms and an alias table of n entries (null: the one
// on the device stays, with `n` its entry count) are consistent with the scene.
static pt_status validate_lights(const pt_scene* s, const PtSceneUniforms* u, const PtSceneAliasEntry* alias, uint32_t n) {
    if (u->emissive_triangles_count > n || n > s->scene_alias_capacity)
        return set_error(PT_ERR_INVALID, "alias table size %u does not fit (capacity %u, emissive %u)", n,
                         s->scene_alias_capacity, u->emissive_triangles_count);
    if (alias)
        for (uint32_t i = 0; i < u->emissive_triangles_count; ++i)
            if (alias[i].alias_index >= u->emissive_triangles_count || alias[i].triangle_index >= s->triangle_count)
                return set_error(PT_ERR_INVALID, "scene alias entry %u out of range", i);
    const float pg = u->light_probabilities[0], ps = u->light_probabilities[1];
    if ((ps > 0.0f || (pg + ps > 0.0f && pg < 1.0f)) && s->skybox_alias_count == 0)
        return set_error(PT_ERR_INVALID, "sky sampling without a sky alias table");
    if (pg > 0.0f && u->emissive_triangles_count == 0)
        return set_error(PT_ERR_INVALID, "geometry light probability without emitters");
    return PT_OK;
}

// Copies n alias entries to the device in stream order (after queued renders).
static pt_status write_alias(pt_scene* s, const PtSceneAliasEntry* alias, uint32_t n) {
    if (!n) return PT_OK;
    hipError_t e = hipMemcpyAsync((void*)s->view.scene_alias, alias, n * sizeof(PtSceneAliasEntry),
                                  hipMemcpyHostToDevice, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "alias update");
}

static void set_skybox_uniforms(pt_scene* s, const PtSkyboxUniforms& sky) {
    PtSkyboxUniforms k = sky;
    k.width = s->view.skybox.width;  // the sky image itself is not replaced
    k.height = s->view.skybox.height;
    s->view.skybox = k;
}

pt_status pt_scene_update_lights(pt_scene* s, const PtSceneUniforms* u, const PtSceneAliasEntry* alias, uint32_t n,
                                 const PtSkyboxUniforms* sky) {
    if (!s || !u || (n && !alias)) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status ist = scene_require_idle(s, "pt_scene_update_lights")) return ist;
    pt_status st = validate_lights(s, u, alias, n);
    if (st) return st;
    if (pt_status dst = use_device(s->ctx->device)) return dst;
    if ((st = write_alias(s, alias, n))) return st;
    s->view.uniforms = *u;
    s->alias_count = n;
    if (sky) set_skybox_uniforms(s, *sky);
    ++s->version;
    return PT_OK;
}

pt_status pt_scene_apply_updates(pt_scene* s, const pt_world_updates* u) {
    if (!s || !u) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status st = scene_require_idle(s, "pt_scene_apply_updates")) return st;
    if (u->material_count && (!u->material_indices || !u->materials))
        return set_error(PT_ERR_INVALID, "material updates without arrays");
    if (u->scene_alias_changed && u->scene_alias_count && !u->scene_alias)
        return set_error(PT_ERR_INVALID, "alias update without entries");
    for (uint32_t i = 0; i < u->material_count; ++i) {
        if (u->material_indices[i] >= s->material_count)
            return set_error(PT_ERR_INVALID, "material index %u out of range", u->material_indices[i]);
        if (u->materials[i].alpha_mode != PT_ALPHA_MODE_OPAQUE && s->view.all_opaque)
            return set_error(PT_ERR_INVALID, "switching a material to alpha-tested needs a new scene");
    }
    const PtSceneUniforms& uni = u->scene_uniforms_changed ? u->scene_uniforms : s->view.uniforms;
    const uint32_t n = u->scene_alias_changed ? u->scene_alias_count : s->alias_count;
    if (u->scene_uniforms_changed || u->scene_alias_changed) {
        pt_status st = validate_lights(s, &uni, u->scene_alias_changed ? u->scene_alias : nullptr, n);
        if (st) return st;
    }
    if (pt_status dst = use_device(s->ctx->device)) return dst;
    if (u->material_count) {
        pt_status st = pt_scene_update_materials(s, u->material_indices, u->materials, u->material_count);
        if (st) return st;
    }
    if (u->scene_alias_changed) {
        pt_status st = write_alias(s, u->scene_alias, u->scene_alias_count);
        if (st) return st;
        s->alias_count = u->scene_alias_count;
    }
    if (u->scene_uniforms_changed) s->view.uniforms = u->scene_uniforms;
    if (u->skybox_uniforms_changed) set_skybox_uniforms(s, u->skybox_uniforms);
    if (u->camera_changed) s->view.camera = u->camera;
    if (u->reset_accumulation) ++s->version;
    return PT_OK;
}

pt_status pt_scene_intersect(pt_scene* s, const float* rays, uint32_t count, int32_t* hit, float* t, float* v, float* w,
                             uint32_t* pairs, uint32_t* tris) {
    if (!s || !rays || !hit || !t || !v || !w) return set_error(PT_ERR_INVALID, "null argument");
    if (count == 0) return PT_OK;
    Context* c = s->ctx;
    if (pt_status dst = use_device(c->device)) return dst;
    uint32_t grid = std::min<uint32_t>((count + 255) / 256, c->cu_count * 8);
    pt_status st = c->ensure_overflow(s->bvh_depth + 1, grid * 256);
    if (st) return st;
    DeviceBuffer buf;
    size_t nb = (size_t)count * (6 * 4 + 4 * 6);
    if ((st = buf.alloc(nb))) return st;
    char* p = (char*)buf.ptr;
    float* drays = (float*)p;
    int32_t* dh = (int32_t*)(p + (size_t)count * 24);
    float* dt = (float*)(dh + count);
    float* dv = dt + count;
    float* dw = dv + count;
    uint32_t* dp = (uint32_t*)(dw + count);
    uint32_t* dtr = dp + count;
    hipError_t e = hipMemcpyAsync(drays, rays, (size_t)count * 24, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_intersect(s->view, drays, count, dh, dt, dv, dw, dp, dtr, c->ovf, grid, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hit, dh, (size_t)count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t, dt, (size_t)count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(v, dv, (size_t)count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w, dw, (size_t)count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && pairs) e = hipMemcpyAsync(pairs, dp, (size_t)count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && tris) e = hipMemcpyAsync(tris, dtr, (size_t)count * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "pt_scene_intersect");
}

pt_status pt_scene_occluded(pt_scene* s, const float* rays, const float* tmax, const int32_t* skip, uint32_t count,
                            uint8_t* occ) {
    if (!s || !rays || !tmax || !skip || !occ) return set_error(PT_ERR_INVALID, "null argument");
    if (count == 0) return PT_OK;
    Context* c = s->ctx;
    if (pt_status dst = use_device(c->device)) return dst;
    uint32_t grid = std::min<uint32_t>((count + 255) / 256, c->cu_count * 8);
    pt_status st = c->ensure_overflow(s->bvh_depth + 1, grid * 256);
    if (st) return st;
    DeviceBuffer buf;
    if ((st = buf.alloc((size_t)count * (24 + 4 + 4 + 1)))) return st;
    char* p = (char*)buf.ptr;
    float* drays = (float*)p;
    float* dt = (float*)(p + (size_t)count * 24);
    int32_t* dsk = (int32_t*)(dt + count);
    uint8_t* docc = (uint8_t*)(dsk + count);
    hipError_t e = hipMemcpyAsync(drays, rays, (size_t)count * 24, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, tmax, (size_t)count * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dsk, skip, (size_t)count * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_occluded(s->view, drays, dt, dsk, count, docc, c->ovf, grid, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(occ, docc, count, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "pt_scene_occluded");
}

}  // extern "C"
