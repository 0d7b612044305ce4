// kernel_on_host.hip - TEST INFRASTRUCTURE (the oracle side; never part of
// the product). Compiles the device path-tracing code of
// path-tracer_amd/csrc/pt_device.h for the host CPU, unchanged, and renders
// tiles with it. Tests compare its output with the C++ oracle
// (trace_oracle.cpp) bit for bit: that pins the kernel's expressions and
// evaluation order against the restatement of trace.slang on the CPU, without
// a GPU; the GPU tests then pin the hardware run against the same oracle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

// the one device builtin the kernel code uses, for the host
__host__ __device__ inline unsigned int pt_host_float_as_uint(float f) {
    unsigned int u;
    std::memcpy(&u, &f, 4);
    return u;
}
#define __float_as_uint pt_host_float_as_uint
#undef __device__
#define __device__ __host__ __attribute__((device))
#include "../path-tracer_amd/csrc/pt_device.h"
#include "../include/pt_api.h"

using namespace ptd;

namespace {
void srgb_table(float lut[256]) {  // IEC 61966-2-1 EOTF, as the library's context builds it
    for (int i = 0; i < 256; ++i) {
        double c = i / 255.0;
        lut[i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
}
}  // namespace

extern "C" int kernel_on_host_render(const pt_scene_desc* d, uint32_t image_w, uint32_t image_h, uint32_t origin_x,
                                     uint32_t origin_y, uint32_t tile_w, uint32_t tile_h, uint32_t max_bounces,
                                     uint32_t sample, uint32_t pair_leaves, float* out_rgb) {
    DevScene s{};
    s.geo = d->geometries;
    s.attr = d->attributes;
    s.mat = d->materials;
    s.bvh = d->bvh_nodes;
    std::vector<float4> tn(2 * (size_t)d->bvh_node_count);
    for (uint32_t i = 0; i < d->bvh_node_count; ++i) {
        const PtBvhNode& n = d->bvh_nodes[i];
        uint32_t code = child_code(i, n.left_first, n.tri_count);
        float cf, tf;
        std::memcpy(&cf, &code, 4);
        std::memcpy(&tf, &n.tri_count, 4);
        tn[2 * (size_t)i] = make_float4(n.aabb_min[0], n.aabb_min[1], n.aabb_min[2], cf);
        tn[2 * (size_t)i + 1] = make_float4(n.aabb_max[0], n.aabb_max[1], n.aabb_max[2], tf);
    }
    s.tnode = tn.data();
    std::vector<uint8_t> opaque(d->triangle_count);
    bool all_opaque = true;
    for (uint32_t i = 0; i < d->triangle_count; ++i) {
        opaque[i] = d->materials[d->attributes[i].material].alpha_mode == PT_ALPHA_MODE_OPAQUE;
        all_opaque = all_opaque && opaque[i];
    }
    float lut[256];
    srgb_table(lut);
    s.srgb = d->srgb_atlases;
    s.linear = d->linear_atlases;
    s.sky = (const float4*)d->skybox_rgba;
    s.sky_alias = d->skybox_alias;
    s.scene_alias = d->scene_alias;
    s.tri_opaque = opaque.data();
    s.srgb_lut = lut;
    s.camera = d->camera;
    s.skybox = d->skybox;
    s.uniforms = d->scene_uniforms;
    s.atlas_size = d->atlas_size;
    s.triangle_count = d->triangle_count;
    s.all_opaque = all_opaque ? 1u : 0u;
    std::vector<uint32_t> lds(4096);
    Stack st;
    st.lds = (lds_u32*)lds.data();  // plain memory in this host build
    st.lds_stride = 1;
    st.lds_depth = (uint32_t)lds.size();
    st.ovf = nullptr;
    st.ovf_stride = 1;
    st.gtid = 0;
    TraversalCounters ctr{0, 0};
    uint32_t closest = 0, shadow = 0;
    for (uint32_t y = 0; y < tile_h; ++y)
        for (uint32_t x = 0; x < tile_w; ++x) {
            // both leaf forms of the traversal: the paired triangle loads of the
            // traversal kernels, and the plain loop of the whole-path kernels
            f3 L = pair_leaves ? trace_path<false, true>(s, origin_x + x, origin_y + y, image_w, image_h, max_bounces,
                                                        sample, st, ctr, closest, shadow)
                               : trace_path<false, false>(s, origin_x + x, origin_y + y, image_w, image_h, max_bounces,
                                                         sample, st, ctr, closest, shadow);
            float* o = &out_rgb[((size_t)y * tile_w + x) * 3];
            o[0] = L.x;
            o[1] = L.y;
            o[2] = L.z;
        }
    return 0;
}
