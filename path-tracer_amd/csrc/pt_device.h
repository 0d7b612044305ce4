// pt_device.h - device functions of the path-tracing kernels (gfx950).
//
// One restatement of the reference's trace kernel (src/render/trace.slang),
// split into pieces that the megakernel and the wavefront kernels both call:
// ray generation, BVH traversal (closest and any hit), surface shading setup,
// light sampling, BSDF evaluation and sampling. Expressions are evaluated in
// the shader's order with fma exactly where it writes fma, so a path gives
// the same bits in every schedule and on the host oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pt_layout.h"
#include "pt_math.h"

namespace ptd {

using ptm::f3;

// ------------------------------------------------------------------ scene view
// Device pointers + uniforms of one scene: bind group 1 of the reference
// (world_resources.rs:249-390) plus the library's derived arrays.
struct DevScene {
    const PtTriangleGeometry* geo;
    const PtTriangleAttribute* attr;
    const PtMaterial* mat;
    const PtBvhNode* bvh;
    // traversal copy of the nodes: node i is tnode[2i] = (aabb_min, code),
    // tnode[2i + 1] = (aabb_max, tri_count), with the code of child_code()
    // precomputed, so a child pair read needs no decoding. The array starts
    // TNODE_LEAD entries into its allocation, which puts every child pair
    // (odd first index) on a 64-byte boundary.
    const float4* tnode;
    const uint8_t* srgb;       // sRGB atlas layers, RGBA8
    const uint8_t* linear;     // linear atlas layers, RGBA8
    const float4* sky;         // normalised sky, RGBA32F
    const PtSkyboxAliasEntry* sky_alias;
    const PtSceneAliasEntry* scene_alias;
    const uint8_t* tri_opaque; // per triangle: 1 if its material is opaque (null when all are)
    const float* srgb_lut;     // 256 entries: sRGB-encoded byte -> linear (Rgba8UnormSrgb decode)
    PtCamera camera;
    PtSkyboxUniforms skybox;
    PtSceneUniforms uniforms;
    uint32_t atlas_size;
    uint32_t triangle_count;
    uint32_t all_opaque;
    uint32_t pad;
};

constexpr size_t TNODE_LEAD = 2;  // float4 entries before node 0 (32 bytes)

// ------------------------------------------------------------------ RNG
struct Rng {
    uint32_t s;
    __device__ __forceinline__ float next() { return ptm::rand_step(s); }
};

// ------------------------------------------------------------------ textures
// nearest, clamp-to-edge, LOD 0 (world_resources.rs:447-450)
__device__ __forceinline__ float4 atlas_fetch(const float* lut, const uint8_t* base, uint32_t size, int layer, float u,
                                              float v, bool srgb) {
    float fx = floorf(u * (float)size);
    float fy = floorf(v * (float)size);
    uint32_t ix = fx < 0.0f ? 0u : (fx > (float)(size - 1u) ? size - 1u : (uint32_t)fx);
    uint32_t iy = fy < 0.0f ? 0u : (fy > (float)(size - 1u) ? size - 1u : (uint32_t)fy);
    uint32_t texel = *(const uint32_t*)(base + (((uint64_t)layer * size + iy) * size + ix) * 4u);
    uint32_t r = texel & 255u, g = (texel >> 8) & 255u, b = (texel >> 16) & 255u, a = texel >> 24;
    float4 o;
    if (srgb) {
        o.x = lut[r];
        o.y = lut[g];
        o.z = lut[b];
    } else {
        o.x = (float)r / 255.0f;
        o.y = (float)g / 255.0f;
        o.z = (float)b / 255.0f;
    }
    o.w = (float)a / 255.0f;
    return o;
}
__device__ __forceinline__ float4 tex_uv(const float* lut, const uint8_t* base, uint32_t size, int layer,
                                         const float* uvo, float u, float v, bool srgb) {
    return atlas_fetch(lut, base, size, layer, ptm::fma_(u, uvo[2], uvo[0]), ptm::fma_(v, uvo[3], uvo[1]), srgb);
}

// ------------------------------------------------------------------ sky (trace.slang:467-489)
__device__ __forceinline__ void skybox_uv(f3 d, float& u, float& v) {
    u = ptm::frac_(ptm::fma_(ptm::atan2_(d.z, d.x), ptm::INV_2PI, 0.5f));
    v = ptm::atan2_(sqrtf(d.x * d.x + d.z * d.z), d.y) * ptm::INV_PI;
}
// bilinear, clamp-to-edge (world_resources.rs:451-456)
__device__ __forceinline__ f3 sky_bilinear(const DevScene& s, float u, float v) {
    const uint32_t W = s.skybox.width, H = s.skybox.height;
    float x = u * (float)W - 0.5f;
    float y = v * (float)H - 0.5f;
    float x0 = floorf(x), y0 = floorf(y);
    float fx = x - x0, fy = y - y0;
    int ix = (int)x0, iy = (int)y0;
    uint32_t xa = ix < 0 ? 0u : ((uint32_t)ix >= W ? W - 1u : (uint32_t)ix);
    uint32_t xb = ix + 1 < 0 ? 0u : ((uint32_t)(ix + 1) >= W ? W - 1u : (uint32_t)(ix + 1));
    uint32_t ya = iy < 0 ? 0u : ((uint32_t)iy >= H ? H - 1u : (uint32_t)iy);
    uint32_t yb = iy + 1 < 0 ? 0u : ((uint32_t)(iy + 1) >= H ? H - 1u : (uint32_t)(iy + 1));
    float4 p00 = s.sky[(size_t)ya * W + xa], p10 = s.sky[(size_t)ya * W + xb];
    float4 p01 = s.sky[(size_t)yb * W + xa], p11 = s.sky[(size_t)yb * W + xb];
    f3 top = ptm::lerp3(f3{p00.x, p00.y, p00.z}, f3{p10.x, p10.y, p10.z}, fx);
    f3 bot = ptm::lerp3(f3{p01.x, p01.y, p01.z}, f3{p11.x, p11.y, p11.z}, fx);
    return ptm::lerp3(top, bot, fy);
}
__device__ __forceinline__ f3 evaluate_skybox(const DevScene& s, f3 d) {
    float u, v;
    skybox_uv(d, u, v);
    return sky_bilinear(s, u, v) * s.skybox.average_luminance;
}
__device__ __forceinline__ float skybox_texel_solid_angle(const DevScene& s, uint32_t row) {
    float step = ptm::PI / (float)s.skybox.height;
    uint32_t pole_row = min(row, s.skybox.height - 1u - row);
    return (2.0f * ptm::PI / (float)s.skybox.width) * 2.0f * ptm::sin_(((float)pole_row + 0.5f) * step) *
           ptm::sin_(0.5f * step);
}
__device__ __forceinline__ float skybox_light_pdf(const DevScene& s, f3 d, float category_probability) {
    if (category_probability <= 0.0f) return 0.0f;
    float u, v;
    skybox_uv(d, u, v);
    uint32_t px = min((uint32_t)(u * (float)s.skybox.width), s.skybox.width - 1u);
    uint32_t py = min((uint32_t)(v * (float)s.skybox.height), s.skybox.height - 1u);
    return category_probability * s.sky_alias[py * s.skybox.width + px].selection_probability /
           skybox_texel_solid_angle(s, py);
}

// ------------------------------------------------------------------ materials (trace.slang:426-466)
__device__ __forceinline__ void bary_uv(const PtTriangleAttribute& at, float u, float v, float w, float& ou, float& ov) {
    ou = ptm::frac_(u * at.uv[0] + v * at.uv[2] + w * at.uv[4]);
    ov = ptm::frac_(u * at.uv[1] + v * at.uv[3] + w * at.uv[5]);
}
__device__ __forceinline__ float coverage(const DevScene& s, const PtMaterial& m, float u, float v) {
    if (m.alpha_mode == PT_ALPHA_MODE_OPAQUE) return 1.0f;
    float alpha = m.base_color[3];
    if (m.base_color_tex_layer >= 0) alpha *= tex_uv(s.srgb_lut, s.srgb, s.atlas_size, m.base_color_tex_layer, m.base_color_uv, u, v, true).w;
    if (m.alpha_mode == PT_ALPHA_MODE_MASK) return alpha >= m.alpha_cutoff ? 1.0f : 0.0f;
    return alpha;
}
__device__ __forceinline__ bool passes_through(const DevScene& s, uint32_t tri, float v, float w, Rng& rng) {
    if (s.all_opaque || s.tri_opaque[tri]) return false;
    const PtTriangleAttribute& at = s.attr[tri];
    const PtMaterial& m = s.mat[at.material];
    if (m.alpha_mode == PT_ALPHA_MODE_OPAQUE) return false;
    float uu, vv;
    bary_uv(at, 1.0f - v - w, v, w, uu, vv);
    return coverage(s, m, uu, vv) <= rng.next();
}
__device__ __forceinline__ f3 evaluate_emission(const DevScene& s, const PtMaterial& m, float u, float v) {
    f3 color = {m.emission_color_and_luminance[0], m.emission_color_and_luminance[1], m.emission_color_and_luminance[2]};
    if (m.emission_tex_layer >= 0) {
        float4 t = tex_uv(s.srgb_lut, s.srgb, s.atlas_size, m.emission_tex_layer, m.emission_uv, u, v, true);
        color = color * f3{t.x, t.y, t.z};
    }
    return m.emission_weight * m.emission_color_and_luminance[3] * color;
}

// ------------------------------------------------------------------ lights (trace.slang:490-592)
__device__ __forceinline__ f3 vtx(const PtTriangleGeometry& g, int k) { return f3{g.v[3 * k], g.v[3 * k + 1], g.v[3 * k + 2]}; }

__device__ __forceinline__ float triangle_selection_probability(const DevScene& s, uint32_t tri) {
    uint32_t first = 0, end = s.uniforms.emissive_triangles_count;
    while (first < end) {
        uint32_t middle = first + (end - first) / 2u;
        PtSceneAliasEntry e = s.scene_alias[middle];
        if (e.triangle_index < tri) first = middle + 1u;
        else if (e.triangle_index > tri) end = middle;
        else return e.selection_probability;
    }
    return 0.0f;
}
__device__ __forceinline__ float triangle_light_pdf(const DevScene& s, uint32_t tri, f3 dir, float d2, float prob) {
    if (prob <= 0.0f) return 0.0f;
    PtTriangleGeometry g = s.geo[tri];
    f3 v0 = vtx(g, 0), v1 = vtx(g, 1), v2 = vtx(g, 2);
    float projected_area = 0.5f * fabsf(ptm::dot(ptm::cross(v1 - v0, v2 - v0), dir));
    if (projected_area <= 0.0f) return 0.0f;
    return prob * d2 / projected_area;
}
__device__ __forceinline__ float power_heuristic(float pdf, float other) {
    float scale = fmaxf(pdf, other);
    if (scale <= 0.0f) return 0.0f;
    float p = pdf / scale;
    float q = other / scale;
    return p * p / (p * p + q * q);
}

struct LightSample {
    f3 direction;
    float distance;
    f3 emission;
    float pdf;
    int32_t triangle_index;
};
__device__ __forceinline__ LightSample sample_light(const DevScene& s, f3 origin, Rng& rng) {
    LightSample r;
    r.direction = f3{0.0f, 0.0f, 0.0f};
    r.distance = 0.0f;
    r.emission = f3{0.0f, 0.0f, 0.0f};
    r.pdf = 0.0f;
    r.triangle_index = -1;
    const float px = s.uniforms.light_probabilities[0], py = s.uniforms.light_probabilities[1];
    if (px + py <= 0.0f) return r;
    if (rng.next() < px) {
        const uint32_t count = s.uniforms.emissive_triangles_count;
        uint32_t slot = min((uint32_t)(rng.next() * (float)count), count - 1u);
        PtSceneAliasEntry cand = s.scene_alias[slot];
        PtSceneAliasEntry entry = s.scene_alias[rng.next() < cand.threshold ? slot : cand.alias_index];
        PtTriangleGeometry g = s.geo[entry.triangle_index];
        f3 v0 = vtx(g, 0), v1 = vtx(g, 1), v2 = vtx(g, 2);
        float root = sqrtf(rng.next());
        float u = 1.0f - root;
        float v = root * (1.0f - rng.next());
        float w = 1.0f - u - v;
        f3 delta = u * v0 + v * v1 + w * v2 - origin;
        float d2 = ptm::dot(delta, delta);
        if (d2 <= 0.0f) return r;
        r.distance = sqrtf(d2);
        r.direction = delta / r.distance;
        r.triangle_index = (int32_t)entry.triangle_index;
        r.pdf = triangle_light_pdf(s, entry.triangle_index, r.direction, d2, px * entry.selection_probability);
        const PtTriangleAttribute& at = s.attr[entry.triangle_index];
        float uu, vv;
        bary_uv(at, u, v, w, uu, vv);
        const PtMaterial& m = s.mat[at.material];
        r.emission = coverage(s, m, uu, vv) * evaluate_emission(s, m, uu, vv);
    } else {
        const uint32_t W = s.skybox.width, H = s.skybox.height;
        const uint32_t count = W * H;
        uint32_t slot = min((uint32_t)(rng.next() * (float)count), count - 1u);
        PtSkyboxAliasEntry cand = s.sky_alias[slot];
        uint32_t index = rng.next() < cand.threshold ? slot : cand.alias_index;
        uint32_t row = index / W, column = index % W;
        float phi = (((float)column + rng.next()) / (float)W - 0.5f) * (2.0f * ptm::PI);
        bool north = row < H / 2u;
        uint32_t pole_row = north ? row : H - 1u - row;
        float half_step = 0.5f * ptm::PI / (float)H;
        float a = ptm::sin_((float)pole_row * half_step);
        float b = ptm::sin_((float)(pole_row + 1u) * half_step);
        float shs = ptm::lerp_(a * a, b * b, rng.next());
        float cos_theta = 1.0f - 2.0f * shs;
        float sin_theta = 2.0f * sqrtf(fmaxf(0.0f, shs * (1.0f - shs)));
        float sn, cs;
        ptm::sincos_(phi, sn, cs);
        r.direction = f3{sin_theta * cs, north ? cos_theta : -cos_theta, sin_theta * sn};
        r.distance = ptm::FLT_MAX_;
        r.pdf = skybox_light_pdf(s, r.direction, py);
        r.emission = evaluate_skybox(s, r.direction);
    }
    return r;
}

// ------------------------------------------------------------------ traversal (trace.slang:597-884)
struct Watertight {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
__device__ __forceinline__ Watertight watertight_setup(f3 d, f3 rcp) {
    f3 ad = ptm::abs3(d);
    int kxy = ad.x > ad.y ? 0 : 1;
    int kz = ptm::comp(ad, kxy) > ad.z ? kxy : 2;
    int kx_val = kz == 2 ? 0 : kz + 1;
    int ky_val = kx_val == 2 ? 0 : kx_val + 1;
    bool is_neg = ptm::comp(d, kz) < 0.0f;
    Watertight w;
    w.kx = is_neg ? ky_val : kx_val;
    w.ky = is_neg ? kx_val : ky_val;
    w.kz = kz;
    w.Sz = ptm::comp(rcp, kz);
    w.Sx = -ptm::comp(d, w.kx) * w.Sz;
    w.Sy = -ptm::comp(d, w.ky) * w.Sz;
    return w;
}

// Triangle test of trace.slang:805-835, after the ray-space permutation has
// been applied to the origin (o_p = origin permuted as (kx, ky, kz)).
__device__ __forceinline__ bool tri_test(const float* __restrict__ v, f3 o, const Watertight& w, float hit_t,
                                         float& tt, float& bv, float& bw) {
    f3 A = f3{v[0], v[1], v[2]} - o, B = f3{v[3], v[4], v[5]} - o, C = f3{v[6], v[7], v[8]} - o;
    float Akz = ptm::comp(A, w.kz), Bkz = ptm::comp(B, w.kz), Ckz = ptm::comp(C, w.kz);
    float Ax = ptm::fma_(w.Sx, Akz, ptm::comp(A, w.kx));
    float Ay = ptm::fma_(w.Sy, Akz, ptm::comp(A, w.ky));
    float Bx = ptm::fma_(w.Sx, Bkz, ptm::comp(B, w.kx));
    float By = ptm::fma_(w.Sy, Bkz, ptm::comp(B, w.ky));
    float Cx = ptm::fma_(w.Sx, Ckz, ptm::comp(C, w.kx));
    float Cy = ptm::fma_(w.Sy, Ckz, ptm::comp(C, w.ky));
    float U = ptm::fma_(Cx, By, -Cy * Bx);
    float V = ptm::fma_(Ax, Cy, -Ay * Cx);
    float W = ptm::fma_(Bx, Ay, -By * Ax);
    if (fminf(U, fminf(V, W)) < 0.0f && fmaxf(U, fmaxf(V, W)) > 0.0f) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float inv_det = 1.0f / det;
    float t = w.Sz * ptm::fma_(U, Akz, ptm::fma_(V, Bkz, W * Ckz)) * inv_det;
    if (!(t > 0.0f && t < hit_t)) return false;
    tt = t;
    bv = V * inv_det;
    bw = W * inv_det;
    return true;
}

__device__ __forceinline__ float intersect_aabb(float4 lo, float4 hi, f3 neg, f3 rcp, float hit_t) {
    float t0x = ptm::fma_(lo.x, rcp.x, neg.x), t0y = ptm::fma_(lo.y, rcp.y, neg.y), t0z = ptm::fma_(lo.z, rcp.z, neg.z);
    float t1x = ptm::fma_(hi.x, rcp.x, neg.x), t1y = ptm::fma_(hi.y, rcp.y, neg.y), t1z = ptm::fma_(hi.z, rcp.z, neg.z);
    float t_near = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    float t_far = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    return (t_near <= t_far && t_far > 0.0f && t_near < hit_t) ? fmaxf(0.0f, t_near) : -1.0f;
}

// Stack entries and the "current node" use one 32-bit code:
//   bit31 = 1          leaf with 1 or 2 triangles: start = (e >> 1) & 0x3fffffff, count = (e & 1) + 1
//   bit31 = 0, bit30=1 leaf with more triangles: node index = e & 0x3fffffff (re-read the node)
//   bit31 = 0, bit30=0 internal node: e = left_first (its children are the pair e, e + 1)
// so a child's box pair is fetched once (64 contiguous bytes) and no node is
// ever re-read on pop, unlike trace.slang:793 which re-reads every node.
constexpr uint32_t CODE_SMALL_LEAF = 0x80000000u;
constexpr uint32_t CODE_BIG_LEAF = 0x40000000u;
__host__ __device__ __forceinline__ uint32_t child_code(uint32_t node_index, uint32_t left_first, uint32_t tri_count) {
    if (tri_count == 0u) return left_first;
    if (tri_count <= 2u) return CODE_SMALL_LEAF | (left_first << 1) | (tri_count - 1u);
    return CODE_BIG_LEAF | node_index;
}

// Per-thread traversal stack: `lds_depth` entries in LDS (lane-interleaved,
// conflict-free), the rest in a grid-strided global overflow area. In the
// device build the LDS part is typed as LDS (address space 3), so its reads
// and writes are ds_read/ds_write and never merge with the overflow accesses
// into flat (either-space) instructions.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint32_t lds_u32;
#else
typedef uint32_t lds_u32;
#endif
__device__ __forceinline__ lds_u32* as_lds(uint32_t* p) { return (lds_u32*)p; }
struct Stack {
    lds_u32* lds;         // lds[k * stride + lane]
    uint32_t* ovf;        // ovf[(k - lds_depth) * ovf_stride + gtid]
    uint32_t lds_stride;
    uint32_t lds_depth;
    uint32_t ovf_stride;
    uint32_t gtid;
    uint32_t sp;
    __device__ __forceinline__ void push(uint32_t e) {
        if (sp < lds_depth) lds[sp * lds_stride] = e;
        else ovf[(size_t)(sp - lds_depth) * ovf_stride + gtid] = e;
        ++sp;
    }
    // The LDS read is unconditional (clamped index) and the overflow read is a
    // separate, rarely taken branch: a pop reads LDS with a ds_read and waits
    // only for LDS, instead of one flat load through a selected address space.
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        uint32_t e = lds[(sp < lds_depth ? sp : lds_depth - 1u) * lds_stride];
        if (sp >= lds_depth) e = ovf[(size_t)(sp - lds_depth) * ovf_stride + gtid];
        return e;
    }
};

struct TraversalCounters {
    uint32_t pairs, tris;
};

// Leaf step of the closest-hit walk: the triangles of leaf `cur`, strictly
// nearer hits accepted, alpha-tested candidates drawing as passes_through does.
// PAIR: a leaf of two triangles requests both before the first test, then
// tests them in order.
template <bool COUNT, bool PAIR>
__device__ __forceinline__ void leaf_closest(const DevScene& s, uint32_t cur, f3 o, const Watertight& wt, Rng& rng,
                                             float& hit_t, int32_t& hit, float& hit_v, float& hit_w,
                                             TraversalCounters& ctr) {
    if (PAIR && (cur & CODE_SMALL_LEAF)) {
        const uint32_t t0 = (cur >> 1) & 0x3fffffffu, t1 = t0 + 1u;
        const PtTriangleGeometry g0 = s.geo[t0];
        PtTriangleGeometry g1;
        if (cur & 1u) g1 = s.geo[t1];
        float tt, bv, bw;
        if (COUNT) ++ctr.tris;
        if (tri_test(g0.v, o, wt, hit_t, tt, bv, bw) && !passes_through(s, t0, bv, bw, rng)) {
            hit_t = tt;
            hit = (int32_t)t0;
            hit_v = bv;
            hit_w = bw;
        }
        if (cur & 1u) {
            if (COUNT) ++ctr.tris;
            if (tri_test(g1.v, o, wt, hit_t, tt, bv, bw) && !passes_through(s, t1, bv, bw, rng)) {
                hit_t = tt;
                hit = (int32_t)t1;
                hit_v = bv;
                hit_w = bw;
            }
        }
        return;
    }
    uint32_t first, count;
    if (cur & CODE_SMALL_LEAF) {
        first = (cur >> 1) & 0x3fffffffu;
        count = (cur & 1u) + 1u;
    } else {
        PtBvhNode n = s.bvh[cur & 0x3fffffffu];
        first = n.left_first;
        count = n.tri_count;
    }
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t ti = first + i;
        if (COUNT) ++ctr.tris;
        float tt, bv, bw;
        if (tri_test(s.geo[ti].v, o, wt, hit_t, tt, bv, bw)) {
            if (passes_through(s, ti, bv, bw, rng)) continue;
            hit_t = tt;
            hit = (int32_t)ti;
            hit_v = bv;
            hit_w = bw;
        }
    }
}

// Leaf step of the any-hit walk: true when a triangle of leaf `cur` other
// than light_tri blocks the ray (an alpha-tested one unless it passes through).
template <bool COUNT, bool PAIR>
__device__ __forceinline__ bool leaf_any(const DevScene& s, uint32_t cur, f3 o, const Watertight& wt, float hit_t,
                                         int32_t light_tri, Rng& rng, TraversalCounters& ctr) {
    if (PAIR && (cur & CODE_SMALL_LEAF)) {
        // as in the closest-hit leaf: both triangles requested, then tested in order
        const uint32_t t0 = (cur >> 1) & 0x3fffffffu, t1 = t0 + 1u;
        const PtTriangleGeometry g0 = s.geo[t0];
        PtTriangleGeometry g1;
        if (cur & 1u) g1 = s.geo[t1];
        float tt, bv, bw;
        if ((int32_t)t0 != light_tri) {
            if (COUNT) ++ctr.tris;
            if (tri_test(g0.v, o, wt, hit_t, tt, bv, bw) && !passes_through(s, t0, bv, bw, rng)) return true;
        }
        if ((cur & 1u) && (int32_t)t1 != light_tri) {
            if (COUNT) ++ctr.tris;
            if (tri_test(g1.v, o, wt, hit_t, tt, bv, bw) && !passes_through(s, t1, bv, bw, rng)) return true;
        }
        return false;
    }
    uint32_t first, count;
    if (cur & CODE_SMALL_LEAF) {
        first = (cur >> 1) & 0x3fffffffu;
        count = (cur & 1u) + 1u;
    } else {
        PtBvhNode n = s.bvh[cur & 0x3fffffffu];
        first = n.left_first;
        count = n.tri_count;
    }
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t ti = first + i;
        if ((int32_t)ti == light_tri) continue;
        if (COUNT) ++ctr.tris;
        float tt, bv, bw;
        if (tri_test(s.geo[ti].v, o, wt, hit_t, tt, bv, bw) && !passes_through(s, ti, bv, bw, rng)) return true;
    }
    return false;
}

// Closest hit (trace.slang:785-884). Visits children nearer-first with ties
// to the right child, accepts strictly nearer hits, so the hit index equals
// the reference's on the same BVH. rng is advanced by alpha-tested
// candidates exactly as passes_through does.
// PAIR: a leaf of one or two triangles loads both before testing them (in
// order). It pays in the traversal-only kernels; in the register-heavy
// whole-path kernels (megakernel, inline shadow rays of the shade kernel) the
// plain loop is faster.
template <bool COUNT, bool PAIR = true>
__device__ __forceinline__ int32_t closest_hit(const DevScene& s, f3 o, f3 d, f3 rcp, Rng& rng, Stack& st,
                                               float& hit_t, float& hit_v, float& hit_w, TraversalCounters& ctr) {
    const Watertight wt = watertight_setup(d, rcp);
    const f3 neg = (-o) * rcp;
    hit_t = ptm::FLT_MAX_;
    hit_v = 0.0f;
    hit_w = 0.0f;
    int32_t hit = -1;
    const float4* __restrict__ nodes = s.tnode;
    uint32_t cur = __float_as_uint(nodes[0].w);
    st.sp = 0;
    for (;;) {
        if ((cur & (CODE_SMALL_LEAF | CODE_BIG_LEAF)) == 0u) {
            // internal: test the child pair at cur, cur + 1
            float4 l0 = nodes[2u * cur], l1 = nodes[2u * cur + 1u];
            float4 r0 = nodes[2u * cur + 2u], r1 = nodes[2u * cur + 3u];
            if (COUNT) ++ctr.pairs;
            float dl = intersect_aabb(l0, l1, neg, rcp, hit_t);
            float dr = intersect_aabb(r0, r1, neg, rcp, hit_t);
            bool hl = dl >= 0.0f, hr = dr >= 0.0f;
            const uint32_t cl = __float_as_uint(l0.w), cr = __float_as_uint(r0.w);
            if (hl && hr) {
                bool left_closer = dl < dr;
                st.push(left_closer ? cr : cl);
                cur = left_closer ? cl : cr;
                continue;
            } else if (hl) {
                cur = cl;
                continue;
            } else if (hr) {
                cur = cr;
                continue;
            }
        } else {
            leaf_closest<COUNT, PAIR>(s, cur, o, wt, rng, hit_t, hit, hit_v, hit_w, ctr);
        }
        if (st.sp == 0u) break;
        cur = st.pop();
    }
    return hit;
}

// Any hit (trace.slang:614-717), skipping the sampled light triangle.
template <bool COUNT, bool PAIR = true>
__device__ __forceinline__ bool occluded(const DevScene& s, f3 o, f3 d, float hit_t, int32_t light_tri, Rng& rng,
                                         Stack& st, TraversalCounters& ctr) {
    const f3 rcp = f3{1.0f, 1.0f, 1.0f} / d;
    const Watertight wt = watertight_setup(d, rcp);
    const f3 neg = (-o) * rcp;
    const float4* __restrict__ nodes = s.tnode;
    uint32_t cur = __float_as_uint(nodes[0].w);
    st.sp = 0;
    for (;;) {
        if ((cur & (CODE_SMALL_LEAF | CODE_BIG_LEAF)) == 0u) {
            float4 l0 = nodes[2u * cur], l1 = nodes[2u * cur + 1u];
            float4 r0 = nodes[2u * cur + 2u], r1 = nodes[2u * cur + 3u];
            if (COUNT) ++ctr.pairs;
            float dl = intersect_aabb(l0, l1, neg, rcp, hit_t);
            float dr = intersect_aabb(r0, r1, neg, rcp, hit_t);
            bool hl = dl >= 0.0f, hr = dr >= 0.0f;
            const uint32_t cl = __float_as_uint(l0.w), cr = __float_as_uint(r0.w);
            if (hl && hr) {
                bool left_closer = dl < dr;
                st.push(left_closer ? cr : cl);
                cur = left_closer ? cl : cr;
                continue;
            } else if (hl) {
                cur = cl;
                continue;
            } else if (hr) {
                cur = cr;
                continue;
            }
        } else if (leaf_any<COUNT, PAIR>(s, cur, o, wt, hit_t, light_tri, rng, ctr)) {
            st.sp = 0;
            return true;
        }
        if (st.sp == 0u) break;
        cur = st.pop();
    }
    return false;
}

// ------------------------------------------------------------------ step-wise traversal
// The same closest-hit / any-hit walks as above, cut into single steps so a
// lane can finish its ray and start the next one while the other lanes of
// its wave are still traversing (no wave-wide wait for the slowest ray).
// Visiting order, tests and acceptance are exactly those of closest_hit and
// occluded, so results are identical.
struct RayState {
    f3 o, rcp, neg;
    Watertight wt;
    float hit_t, hit_v, hit_w;
    int32_t hit;       // closest: triangle index; shadow: 1 once occluded
    int32_t light;     // shadow: light triangle to skip
    uint32_t cur;      // current node code
    bool shadow;
};

__device__ __forceinline__ void ray_begin(const DevScene& s, RayState& r, f3 o, f3 d, bool shadow, float tmax,
                                          int32_t light, Stack& st) {
    r.o = o;
    r.rcp = f3{1.0f, 1.0f, 1.0f} / d;
    r.neg = (-o) * r.rcp;
    r.wt = watertight_setup(d, r.rcp);
    r.hit_t = shadow ? tmax : ptm::FLT_MAX_;
    r.hit_v = r.hit_w = 0.0f;
    r.hit = shadow ? 0 : -1;
    r.light = light;
    r.shadow = shadow;
    r.cur = __float_as_uint(s.tnode[0].w);
    st.sp = 0;
}

// One traversal step; returns true when the ray is finished.
template <bool COUNT>
__device__ __forceinline__ bool ray_step(const DevScene& s, RayState& r, Rng& rng, Stack& st, TraversalCounters& ctr) {
    const float4* __restrict__ nodes = s.tnode;
    if ((r.cur & (CODE_SMALL_LEAF | CODE_BIG_LEAF)) == 0u) {
        const uint32_t cur = r.cur;
        float4 l0 = nodes[2u * cur], l1 = nodes[2u * cur + 1u];
        float4 r0 = nodes[2u * cur + 2u], r1 = nodes[2u * cur + 3u];
        if (COUNT) ++ctr.pairs;
        float dl = intersect_aabb(l0, l1, r.neg, r.rcp, r.hit_t);
        float dr = intersect_aabb(r0, r1, r.neg, r.rcp, r.hit_t);
        bool hl = dl >= 0.0f, hr = dr >= 0.0f;
        const uint32_t cl = __float_as_uint(l0.w), cr = __float_as_uint(r0.w);
        if (hl && hr) {
            bool left_closer = dl < dr;
            st.push(left_closer ? cr : cl);
            r.cur = left_closer ? cl : cr;
            return false;
        }
        if (hl || hr) {
            r.cur = hl ? cl : cr;
            return false;
        }
    } else {
        uint32_t first, count;
        if (r.cur & CODE_SMALL_LEAF) {
            first = (r.cur >> 1) & 0x3fffffffu;
            count = (r.cur & 1u) + 1u;
        } else {
            PtBvhNode n = s.bvh[r.cur & 0x3fffffffu];
            first = n.left_first;
            count = n.tri_count;
        }
        for (uint32_t i = 0; i < count; ++i) {
            const uint32_t ti = first + i;
            if (r.shadow && (int32_t)ti == r.light) continue;
            if (COUNT) ++ctr.tris;
            float tt, bv, bw;
            if (!tri_test(s.geo[ti].v, r.o, r.wt, r.hit_t, tt, bv, bw)) continue;
            if (r.shadow) {
                if (!passes_through(s, ti, bv, bw, rng)) {
                    r.hit = 1;
                    return true;
                }
            } else {
                if (passes_through(s, ti, bv, bw, rng)) continue;
                r.hit_t = tt;
                r.hit = (int32_t)ti;
                r.hit_v = bv;
                r.hit_w = bw;
            }
        }
    }
    if (st.sp == 0u) return true;
    r.cur = st.pop();
    return false;
}

// One traversal round for the refill schedule: walk internal nodes until a
// leaf is current (the while-while form), test that leaf, pop. Returns true
// when the ray is finished. The node order, tests and acceptance are those of
// closest_hit (SHADOW false) and occluded (SHADOW true), so results are equal.
template <bool SHADOW, bool COUNT>
__device__ __forceinline__ bool ray_advance(const DevScene& s, RayState& r, Rng& rng, Stack& st, TraversalCounters& ctr) {
    const float4* __restrict__ nodes = s.tnode;
    uint32_t cur = r.cur;
    while ((cur & (CODE_SMALL_LEAF | CODE_BIG_LEAF)) == 0u) {
        float4 l0 = nodes[2u * cur], l1 = nodes[2u * cur + 1u];
        float4 r0 = nodes[2u * cur + 2u], r1 = nodes[2u * cur + 3u];
        if (COUNT) ++ctr.pairs;
        float dl = intersect_aabb(l0, l1, r.neg, r.rcp, r.hit_t);
        float dr = intersect_aabb(r0, r1, r.neg, r.rcp, r.hit_t);
        bool hl = dl >= 0.0f, hr = dr >= 0.0f;
        const uint32_t cl = __float_as_uint(l0.w), cr = __float_as_uint(r0.w);
        if (hl && hr) {
            bool left_closer = dl < dr;
            st.push(left_closer ? cr : cl);
            cur = left_closer ? cl : cr;
        } else if (hl || hr) {
            cur = hl ? cl : cr;
        } else {
            if (st.sp == 0u) return true;
            cur = st.pop();
        }
    }
    uint32_t first, count;
    if (cur & CODE_SMALL_LEAF) {
        first = (cur >> 1) & 0x3fffffffu;
        count = (cur & 1u) + 1u;
    } else {
        PtBvhNode n = s.bvh[cur & 0x3fffffffu];
        first = n.left_first;
        count = n.tri_count;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t ti = first + i;
        if (SHADOW && (int32_t)ti == r.light) continue;
        if (COUNT) ++ctr.tris;
        float tt, bv, bw;
        if (!tri_test(s.geo[ti].v, r.o, r.wt, r.hit_t, tt, bv, bw)) continue;
        if (SHADOW) {
            if (!passes_through(s, ti, bv, bw, rng)) {
                r.hit = 1;
                st.sp = 0;
                return true;
            }
        } else {
            if (passes_through(s, ti, bv, bw, rng)) continue;
            r.hit_t = tt;
            r.hit = (int32_t)ti;
            r.hit_v = bv;
            r.hit_w = bw;
        }
    }
    if (st.sp == 0u) return true;
    r.cur = st.pop();
    return false;
}

// ------------------------------------------------------------------ camera ray (trace.slang:736-748)
__device__ __forceinline__ f3 camera_origin(const DevScene& s) {
    return f3{s.camera.position[0], s.camera.position[1], s.camera.position[2]};
}
__device__ __forceinline__ f3 primary_direction(const DevScene& s, uint32_t px, uint32_t py, uint32_t iw, uint32_t ih,
                                                Rng& rng) {
    float r0 = rng.next();
    float r1 = rng.next();
    float u = ((float)px + r0) / (float)iw;
    float v = ((float)py + r1) / (float)ih;
    const PtCamera& c = s.camera;
    f3 llc = {c.lower_left_corner_minus_position[0], c.lower_left_corner_minus_position[1],
              c.lower_left_corner_minus_position[2]};
    f3 h = {c.horizontal[0], c.horizontal[1], c.horizontal[2]};
    f3 vv = {c.vertical[0], c.vertical[1], c.vertical[2]};
    return ptm::normalize(llc + u * h + v * vv);
}

// ------------------------------------------------------------------ BSDF (trace.slang:207-424)
__device__ __forceinline__ f3 orthonormal_tangent(f3 tangent, f3 normal) {
    f3 projected = tangent - ptm::dot(normal, tangent) * normal;
    float length_squared = ptm::dot(projected, projected);
    if (length_squared > 1e-12f) return projected * ptm::rsqrt_(length_squared);
    float sg = normal.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sg + normal.z);
    return f3{ptm::fma_(sg * normal.x * normal.x, a, 1.0f), sg * normal.x * normal.y * a, -sg * normal.x};
}
__device__ __forceinline__ float fifth_power(float x) { return x * x * x * x * x; }
// ior is a modulated IOR (specular_ior_modulated: within [2^-25, 2^25] or NaN)
// and cos_i the cosine between two unit vectors. So the first division's
// numerator 1 - cos_i^2 is +0 or at least 2^-24 in magnitude, and its
// denominator lies in [2^-50, 2^50]: the domain of div_nt_. Below, x < 1 makes
// 1 - x at least 2^-24 (or NaN): the domain of sqrt_nt_.
__device__ __forceinline__ float fresnel_exact(float ior, float cos_i) {
    float x = ptm::div_nt_(ptm::fma_(-cos_i, cos_i, 1.0f), ior * ior);
    if (x >= 1.0f) return 1.0f;
    float cos_t = ptm::sqrt_nt_(1.0f - x);
    float Rs = ptm::fma_(-ior, cos_t, cos_i) / ptm::fma_(ior, cos_t, cos_i);
    float Rp = ptm::fma_(-ior, cos_i, cos_t) / ptm::fma_(ior, cos_i, cos_t);
    return 0.5f * (Rs * Rs + Rp * Rp);
}
// the root's argument is 1 plus a ratio of squares: at least 1, +inf or NaN
__device__ __forceinline__ float ggx_lambda(f3 v, float ax, float ay) {
    return ptm::fma_(0.5f, ptm::sqrt_nt_(1.0f + (ax * ax * v.x * v.x + ay * ay * v.y * v.y) / (v.z * v.z)), -0.5f);
}
// BOUNDED: the caller has checked wo.z >= 0 and ax, ay <= 2^56. The result's
// normalisation then takes the restricted-domain root and division. Its
// squared length is at least h.z^2 >= 2^-48: wo_warped.z >= 0 and r is a
// random float at most 1 - 2^-23, so z + wo_warped.z, which is
// (1 - r)(1 + wo_warped.z) >= 2^-23 less the rounding of z (at most 2^-25,
// or 2^-24 when 1 + wo_warped.z > 2), is at least 3 * 2^-25. It is at most
// 2^118, as |h.x|, |h.y| <= 2 (plus rounding) and ax, ay <= 2^56. So the root
// lies in [2^-24, 2^59] and 1 / root in div_nt_'s domain.
template <bool BOUNDED = false>
__device__ __forceinline__ f3 sample_vndf(f3 wo, float ax, float ay, float r, float two_pi_r) {
    f3 wo_warped = ptm::normalize(f3{wo.x * ax, wo.y * ay, wo.z});
    float z = ptm::fma_(1.0f - r, 1.0f + wo_warped.z, -wo_warped.z);
    // 1 - z^2 of a float z is 0, at least 2^-24 in magnitude, or NaN; clamped to [0, 1]
    float sin_theta = ptm::sqrt_nt_(ptm::saturate_(ptm::fma_(-z, z, 1.0f)));
    float sn, cs;
    ptm::sincos_(two_pi_r, sn, cs);
    float x = sin_theta * cs;
    float y = sin_theta * sn;
    f3 h = f3{x, y, z} + wo_warped;
    const f3 hw = f3{h.x * ax, h.y * ay, h.z};
    if (BOUNDED) return hw * ptm::div_nt_(1.0f, ptm::sqrt_nt_(ptm::dot(hw, hw)));  // = normalize(hw)
    return ptm::normalize(hw);
}
__device__ __forceinline__ float specular_ior_modulated(float ior, float weight) {
    float num = ior - 1.0f;
    float den = ior + 1.0f;
    float f0 = num * num / (den * den);
    float f0_sqrt = ptm::sign_(ior - 1.0f) * sqrtf(fminf(weight * f0, ptm::bits_to_float(0x3f800000u - 1u)));
    return (1.0f + f0_sqrt) / (1.0f - f0_sqrt);
}
__device__ __forceinline__ float e_fon_exact(float mu, float r) {
    float AF = 1.0f / ptm::fma_(0.28779340921f, r, 1.0f);
    float Si = ptm::sqrt_nt_(ptm::fma_(-mu, mu, 1.0f));  // 1 - mu^2: 0, NaN or at least 2^-24 in magnitude
    float G = Si * (ptm::acos_(mu) - Si * mu) + ptm::bits_to_float(0x3f2aaaabu) * ((Si / mu) * (1.0f - (Si * Si * Si)) - Si);
    return AF + r * AF * ptm::INV_PI * G;
}

// BSDF state of one surface hit (the values trace.slang computes at 1097-1181)
struct Bsdf {
    f3 base_color;
    float base_weight, base_metalness, base_diffuse_roughness;
    f3 specular_color;
    float specular_weight, sim, ax, ay, one_plus_lambda_wo;
    f3 F0_conductor, E_dielectric;
    float ltc_a, ltc_b, ltc_c, ltc_d;
    float bxx, bxy;  // LTC basis row X = (bxx, bxy, 0); Y = (-bxy, bxx, 0); Z = (0, 0, 1)
    float P_u, P_c, p_ggx, p_diffuse;
    // evaluate_lobes' factors that depend on wo and the material only
    // (lobe_wo_terms), computed once per hit for both of its calls
    float G1;        // 1 / (1 + Lambda(wo))
    f3 ss_scale;     // rho / pi * A_F of the single-scattering EON lobe
    f3 ms_wo;        // rho_ms / pi * max(1e-7, 1 - E_F(mu_o)) of the multiple-scattering lobe
    float ms_den;    // max(1e-7, 1 - E_F average)
    float detM_sq;   // det(M)^2 of the LTC
    float s_pdf;     // the LTC clipping split, also the sampling lerp weight
};

// The factors of evaluate_lobes (trace.slang:229-315) that depend only on wo
// and the material, in the shader's expressions: the shader evaluates them in
// each of its two lobe evaluations per hit, here they are kept in the Bsdf.
__device__ __forceinline__ void lobe_wo_terms(Bsdf& b, f3 wo) {
    const f3 rho = b.base_color;
    const float r = b.base_diffuse_roughness;
    b.G1 = 1.0f / b.one_plus_lambda_wo;
    float AF = 1.0f / ptm::fma_(0.28779340921f, r, 1.0f);
    b.ss_scale = rho * ptm::INV_PI * AF;
    float EFo = e_fon_exact(wo.z, r);
    float avgEF = AF * ptm::fma_(0.07248821245f, r, 1.0f);
    f3 rho_ms = rho * rho * avgEF / (1.0f - rho * (1.0f - avgEF));
    b.ms_wo = (rho_ms * ptm::INV_PI) * fmaxf(1.0e-7f, 1.0f - EFo);
    b.ms_den = fmaxf(1.0e-7f, 1.0f - avgEF);
    const float a = b.ltc_a, bb = b.ltc_b, c = b.ltc_c, d = b.ltc_d;
    float detM = c * ptm::fma_(-bb, d, a);
    b.detM_sq = detM * detM;
    b.s_pdf = ptm::fma_(0.5f, ptm::rsqrt_(ptm::fma_(d, d, 1.0f)), 0.5f);
}

__device__ __forceinline__ void evaluate_lobes(const Bsdf& b, f3 wo, f3 wi, f3& f_total, float& pdf_total) {
    f3 H = ptm::normalize(wo + wi);
    float HdotW = ptm::dot(H, wo);
    const float G1 = b.G1;
    float G2 = 1.0f / (b.one_plus_lambda_wo + ggx_lambda(wi, b.ax, b.ay));
    float D_helper = (H.x * H.x) / (b.ax * b.ax) + (H.y * H.y) / (b.ay * b.ay) + H.z * H.z;
    float D = ptm::INV_PI / (b.ax * b.ay * D_helper * D_helper);
    float ts = D * G2 / (4.0f * wi.z * wo.z);
    float omh = 1.0f - HdotW;
    float k = 0.00105024005f * HdotW * omh;
    f3 fa = (1.0f - b.specular_color) * k;
    f3 fb = b.F0_conductor * 9031.0f + 7776.0f;
    f3 fc = b.F0_conductor - 1.0f;
    f3 F82 = b.F0_conductor - fifth_power(omh) * ptm::fma3(fa, fb, fc);
    f3 f_conductor = ptm::saturate3(b.specular_weight * F82) * ts;
    f3 f_dielectric = b.specular_color * fresnel_exact(b.sim, HdotW) * ts;
    const float r = b.base_diffuse_roughness;
    float s_f = wi.x * wo.x + wi.y * wo.y;
    float sovertF = s_f > 0.0f ? s_f / fmaxf(wo.z, wi.z) : s_f;
    f3 f_ss = b.ss_scale * ptm::fma_(r, sovertF, 1.0f);
    float EFi = e_fon_exact(wi.z, r);
    f3 f_ms = b.ms_wo * fmaxf(1.0e-7f, 1.0f - EFi) / b.ms_den;
    f3 f_diffuse = b.base_weight * (f_ss + f_ms);
    f3 f_glossy_diffuse = f_dielectric + (1.0f - b.E_dielectric) * f_diffuse;
    f_total = ptm::lerp3(f_glossy_diffuse, f_conductor, b.base_metalness);
    float pdf_ggx = D * G1 / (4.0f * wo.z);
    f3 bx = {b.bxx, b.bxy, 0.0f}, by = {-b.bxy, b.bxx, 0.0f}, bz = {0.0f, 0.0f, 1.0f};
    f3 wl = {ptm::dot(bx, wi), ptm::dot(by, wi), ptm::dot(bz, wi)};
    const float a = b.ltc_a, bb = b.ltc_b, c = b.ltc_c, d = b.ltc_d;
    f3 wh = {c * (wl.x - bb * wl.z), (a - bb * d) * wl.y, -c * (d * wl.x - a * wl.z)};
    float lenSqr = ptm::dot(wh, wh);
    float pdf_c = b.detM_sq / (lenSqr * lenSqr) * fmaxf(wh.z, 0.0f) / (ptm::PI * b.s_pdf);
    float pdf_eon = b.P_u * 0.15915494309f + b.P_c * pdf_c;
    pdf_total = b.p_ggx * pdf_ggx + b.p_diffuse * pdf_eon;
}

template <bool BOUNDED>
__device__ __forceinline__ f3 estimate_e_dielectric(f3 wo, float opl, float ax, float ay, float sim, f3 spec_color,
                                                    Rng& rng) {
    float sum_W = 0.0f;
#pragma unroll 2
    for (uint32_t i = 0; i < 16u; ++i) {
        float rand1 = rng.next();
        float two_pi_rand = 2.0f * ptm::PI * rng.next();
        f3 H = sample_vndf<BOUNDED>(wo, ax, ay, rand1, two_pi_rand);
        f3 wi = ptm::reflect(-wo, H);
        if (wi.z <= 0.0f) continue;
        float F = fresnel_exact(sim, ptm::dot(H, wo));
        sum_W += F * opl / (opl + ggx_lambda(wi, ax, ay));
    }
    return spec_color * sum_W / 16.0f;
}

// Shading frame of a hit (trace.slang:929-1091) and BSDF setup (1097-1181).
// Consumes the estimate_e_dielectric draws. Returns the emission and its MIS
// weight separately: the caller adds (throughput * emission) * weight, the
// shader's left-to-right order (trace.slang:991).
struct SurfaceHit {
    f3 hit_point, flat_N, Tt, Bt, N, wo;
};

__device__ __forceinline__ void setup_surface(const DevScene& s, uint32_t hit, float hit_t, float hit_v, float hit_w,
                                              f3 ray_origin, f3 ray_direction, uint32_t bounce,
                                              float previous_bsdf_pdf, Rng& rng, SurfaceHit& sf, Bsdf& b, f3& emission,
                                              float& emission_mis) {
    const PtTriangleAttribute& at = s.attr[hit];
    const float hit_u = 1.0f - hit_v - hit_w;
    float iu, iv;
    bary_uv(at, hit_u, hit_v, hit_w, iu, iv);
    const PtMaterial& mat = s.mat[at.material];
    b.base_color = f3{mat.base_color[0], mat.base_color[1], mat.base_color[2]};
    b.base_weight = mat.base_weight;
    b.base_metalness = mat.base_metalness;
    b.base_diffuse_roughness = mat.base_diffuse_roughness;
    b.specular_color = f3{mat.specular_color_and_weight[0], mat.specular_color_and_weight[1], mat.specular_color_and_weight[2]};
    b.specular_weight = mat.specular_color_and_weight[3];
    const float specular_roughness = mat.specular_roughness;
    const float anisotropy = mat.specular_roughness_anisotropy;
    const float specular_ior = mat.specular_ior;
    if (mat.base_color_tex_layer >= 0) {
        float4 t = tex_uv(s.srgb_lut, s.srgb, s.atlas_size, mat.base_color_tex_layer, mat.base_color_uv, iu, iv, true);
        b.base_color = b.base_color * f3{t.x, t.y, t.z};
    }
    if (mat.metallic_roughness_tex_layer >= 0) {
        float4 t = tex_uv(s.srgb_lut, s.linear, s.atlas_size, mat.metallic_roughness_tex_layer, mat.metallic_roughness_uv, iu, iv, false);
        b.base_metalness *= t.z;
        b.base_diffuse_roughness *= t.y;
    }
    if (mat.specular_color_tex_layer >= 0) {
        float4 t = tex_uv(s.srgb_lut, s.srgb, s.atlas_size, mat.specular_color_tex_layer, mat.specular_color_uv, iu, iv, true);
        b.specular_color = b.specular_color * f3{t.x, t.y, t.z};
    }
    if (mat.specular_weight_tex_layer >= 0) {
        float4 t = tex_uv(s.srgb_lut, s.linear, s.atlas_size, mat.specular_weight_tex_layer, mat.specular_weight_uv, iu, iv, false);
        b.specular_weight *= t.w;
    }
    emission = evaluate_emission(s, mat, iu, iv);
    emission_mis = 1.0f;
    const float prob_geo = s.uniforms.light_probabilities[0];
    if (bounce > 0u && prob_geo > 0.0f && ptm::any_pos(emission)) {
        float light_pdf = triangle_light_pdf(s, hit, ray_direction, hit_t * hit_t, prob_geo * triangle_selection_probability(s, hit));
        emission_mis = power_heuristic(previous_bsdf_pdf, light_pdf);
    }

    sf.hit_point = ray_direction * hit_t + ray_origin;
    const f3 V = -ray_direction;
    PtTriangleGeometry g = s.geo[hit];
    f3 v0 = vtx(g, 0), v1 = vtx(g, 1), v2 = vtx(g, 2);
    f3 flat_N_raw = ptm::normalize(ptm::cross(v1 - v0, v2 - v0));
    float bfm = ptm::dot(V, flat_N_raw) > 0.0f ? 1.0f : -1.0f;
    sf.flat_N = flat_N_raw * bfm;
    f3 n0 = {at.n[0], at.n[1], at.n[2]}, n1 = {at.n[3], at.n[4], at.n[5]}, n2 = {at.n[6], at.n[7], at.n[8]};
    f3 N = ptm::normalize(hit_u * n0 + hit_v * n1 + hit_w * n2) * bfm;
    f3 t0 = {at.t0[0], at.t0[1], at.t0[2]}, t1 = {at.t1[0], at.t1[1], at.t1[2]}, t2 = {at.t2[0], at.t2[1], at.t2[2]};
    const float t0w = at.t0[3];
    f3 Tt = orthonormal_tangent(hit_u * t0 + hit_v * t1 + hit_w * t2, N) * bfm;
    f3 Bt = ptm::cross(N, Tt) * t0w * bfm;
    bool N_changed = false;
    if (mat.normal_tex_layer >= 0) {
        float4 t = tex_uv(s.srgb_lut, s.linear, s.atlas_size, mat.normal_tex_layer, mat.normal_uv, iu, iv, false);
        f3 local_n = f3{t.x, t.y, t.z} * 2.0f - 1.0f;
        local_n = ptm::normalize(f3{local_n.x * mat.normal_scale, local_n.y * mat.normal_scale, local_n.z});
        N = ptm::normalize(local_n.x * Tt + local_n.y * Bt + local_n.z * N);
        N_changed = true;
    }
    float flat_NdotV = ptm::dot(sf.flat_N, V);
    float threshold = 0.5f * flat_NdotV;
    float NdotV = ptm::dot(N, V);
    if (NdotV < threshold) {
        N = ptm::normalize(ptm::lerp3(N, sf.flat_N, (threshold - NdotV) / (flat_NdotV - NdotV)));
        NdotV = ptm::dot(N, V);
        N_changed = true;
    }
    if (N_changed) {
        Tt = orthonormal_tangent(Tt, N);
        Bt = ptm::cross(N, Tt) * t0w * bfm;
    }
    sf.Tt = Tt;
    sf.Bt = Bt;
    sf.N = N;
    // alpha_anisotropic (trace.slang:369-375)
    float isotropy = 1.0f - anisotropy;
    float alpha_t = specular_roughness * specular_roughness * 1.41421356237f * ptm::rsqrt_(ptm::fma_(isotropy, isotropy, 1.0f));
    b.ax = fmaxf(alpha_t, 1e-4f);
    b.ay = fmaxf(alpha_t * isotropy, 1e-4f);
    b.sim = specular_ior_modulated(specular_ior, b.specular_weight);
    b.F0_conductor = b.base_weight * b.base_color;
    const f3 wo = {ptm::dot(Tt, V), ptm::dot(Bt, V), NdotV};
    sf.wo = wo;
    b.one_plus_lambda_wo = 1.0f + ggx_lambda(wo, b.ax, b.ay);
    const float r = b.base_diffuse_roughness, z = wo.z;
    b.ltc_a = 1.0f + r * (0.303392f + (-0.518982f + 0.111709f * z) * z + (-0.276266f + 0.335918f * z) * r);
    b.ltc_b = r * (-1.16407f + 1.15859f * z + (0.150815f - 0.150105f * z) * r) / (z * z * z - 1.43545f);
    b.ltc_c = 1.0f + r * (0.20013f + (-0.506373f + 0.261777f * z) * z);
    b.ltc_d = r * (0.540852f + (-1.01625f + 0.475392f * z) * z) / (-1.0743f + (0.0725628f + z) * z);
    float lsq = wo.x * wo.x + wo.y * wo.y;
    b.bxx = 1.0f;
    b.bxy = 0.0f;
    if (lsq > 0.0f) {
        float inv = ptm::rsqrt_(lsq);
        b.bxx = wo.x * inv;
        b.bxy = wo.y * inv;
    }
    // sample_vndf's bounds: wo in the upper hemisphere (all but rare grazing
    // cases) and any physical roughness; the general form keeps the bits otherwise
    b.E_dielectric = wo.z >= 0.0f && fmaxf(b.ax, b.ay) <= 0x1p56f
                         ? estimate_e_dielectric<true>(wo, b.one_plus_lambda_wo, b.ax, b.ay, b.sim, b.specular_color, rng)
                         : estimate_e_dielectric<false>(wo, b.one_plus_lambda_wo, b.ax, b.ay, b.sim, b.specular_color, rng);
    const f3 LUMA = {0.2126f, 0.7152f, 0.0722f};
    float albedo_gloss = ptm::dot(b.E_dielectric, LUMA);
    float albedo_diffuse = (1.0f - albedo_gloss) * b.base_weight * ptm::dot(b.base_color, LUMA);
    f3 fsm = b.F0_conductor + (1.0f - b.F0_conductor) * fifth_power(1.0f - wo.z);
    float albedo_metal = ptm::dot(fsm, LUMA);
    float w_ggx = ptm::lerp_(albedo_gloss, albedo_metal, b.base_metalness);
    float w_diffuse = (1.0f - b.base_metalness) * albedo_diffuse;
    b.p_ggx = ptm::clamp_(w_ggx / fmaxf(w_ggx + w_diffuse, 1e-7f), 0.1f, 0.9f);
    b.p_diffuse = 1.0f - b.p_ggx;
    b.P_u = ptm::pow_(r, 0.1f) * (0.162925f + (-0.372058f + (0.538233f - 0.290822f * z) * z) * z);
    b.P_c = 1.0f - b.P_u;
    lobe_wo_terms(b, wo);
}

// BSDF direction sample (trace.slang:1212-1260); consumes 3 draws.
__device__ __forceinline__ f3 sample_bsdf(const Bsdf& b, f3 wo, Rng& rng) {
    float two_pi_rand = 2.0f * ptm::PI * rng.next();
    float rand1 = rng.next();
    f3 wi;
    if (rng.next() < b.p_ggx) {
        wi = ptm::reflect(-wo, sample_vndf(wo, b.ax, b.ay, rand1, two_pi_rand));
    } else if (rand1 < b.P_u) {
        float u1 = rand1 / b.P_u;
        float sinTheta = sqrtf(1.0f - u1 * u1);
        float sn, cs;
        ptm::sincos_(two_pi_rand, sn, cs);
        wi = f3{sinTheta * cs, sinTheta * sn, u1};
    } else {
        float u1 = (rand1 - b.P_u) / b.P_c;
        float R = sqrtf(u1);
        float sn, cs;
        ptm::sincos_(two_pi_rand, sn, cs);
        float y = R * sn;
        const float sp = b.s_pdf;
        float x = -ptm::lerp_(sqrtf(ptm::fma_(-y, y, 1.0f)), R * cs, sp);
        f3 wh = {x, y, sqrtf(fmaxf(1.0f - (x * x + y * y), 0.0f))};
        f3 wu = {b.ltc_a * wh.x + b.ltc_b * wh.z, b.ltc_c * wh.y, b.ltc_d * wh.x + wh.z};
        f3 bx = {b.bxx, b.bxy, 0.0f}, by = {-b.bxy, b.bxx, 0.0f}, bz = {0.0f, 0.0f, 1.0f};
        wi = ptm::normalize(wu.x * bx + wu.y * by + wu.z * bz);
    }
    return wi;
}

// ------------------------------------------------------------------ whole path (the megakernel's per-thread loop)
// One path sample (trace.slang:726-1297), returning final_color.
template <bool COUNT, bool PAIR = false>
__device__ __forceinline__ f3 trace_path(const DevScene& s, uint32_t px, uint32_t py, uint32_t iw, uint32_t ih,
                                         uint32_t max_bounces, uint32_t sample_count, Stack& st,
                                         TraversalCounters& ctr, uint32_t& closest_rays, uint32_t& shadow_rays) {
    Rng rng{ptm::rng_seed(px, py, sample_count)};
    f3 ray_origin = {s.camera.position[0], s.camera.position[1], s.camera.position[2]};
    f3 ray_direction = primary_direction(s, px, py, iw, ih, rng);
    f3 rcp = f3{1.0f, 1.0f, 1.0f} / ray_direction;
    f3 throughput = {1.0f, 1.0f, 1.0f};
    f3 final_color = {0.0f, 0.0f, 0.0f};
    float previous_bsdf_pdf = 0.0f;
    const float prob_sky = s.uniforms.light_probabilities[1];
    for (uint32_t bounce = 0; bounce <= max_bounces; ++bounce) {
        float hit_t, hit_v, hit_w;
        if (COUNT) ++closest_rays;
        int32_t hit = closest_hit<COUNT, PAIR>(s, ray_origin, ray_direction, rcp, rng, st, hit_t, hit_v, hit_w, ctr);
        if (hit == -1) {
            float mis = 1.0f;
            if (bounce > 0u) mis = power_heuristic(previous_bsdf_pdf, skybox_light_pdf(s, ray_direction, prob_sky));
            final_color = final_color + throughput * evaluate_skybox(s, ray_direction) * mis;
            break;
        }
        SurfaceHit sf;
        Bsdf b;
        f3 emission;
        float emission_mis;
        setup_surface(s, (uint32_t)hit, hit_t, hit_v, hit_w, ray_origin, ray_direction, bounce, previous_bsdf_pdf, rng,
                      sf, b, emission, emission_mis);
        final_color = final_color + throughput * emission * emission_mis;
        f3 next_origin = sf.hit_point + sf.flat_N * fmaxf(1e-4f, 1e-4f * ptm::length(sf.hit_point));
        LightSample light = sample_light(s, next_origin, rng);
        f3 lwi = {ptm::dot(sf.Tt, light.direction), ptm::dot(sf.Bt, light.direction), ptm::dot(sf.N, light.direction)};
        if (light.pdf > 0.0f && lwi.z > 0.0f && ptm::any_pos(light.emission)) {
            f3 f;
            float pdf;
            evaluate_lobes(b, sf.wo, lwi, f, pdf);
            if (COUNT) ++shadow_rays;
            if (!occluded<COUNT, PAIR>(s, next_origin, light.direction, light.distance, light.triangle_index, rng, st, ctr)) {
                float weight = bounce < max_bounces ? power_heuristic(light.pdf, pdf) : 1.0f;
                final_color = final_color + throughput * f * lwi.z * light.emission * (weight / light.pdf);
            }
        }
        if (bounce == max_bounces) break;
        f3 wi = sample_bsdf(b, sf.wo, rng);
        if (wi.z <= 0.0f) break;
        f3 f;
        float pdf;
        evaluate_lobes(b, sf.wo, wi, f, pdf);
        if (pdf <= 0.0f) break;
        throughput = throughput * (f * wi.z / pdf);
        previous_bsdf_pdf = pdf;
        if (bounce > 3u) {
            float p_survive = ptm::clamp_(fmaxf(throughput.x, fmaxf(throughput.y, throughput.z)), 0.05f, 0.95f);
            if (rng.next() > p_survive) break;
            throughput = throughput / p_survive;
        }
        ray_origin = next_origin;
        ray_direction = wi.x * sf.Tt + wi.y * sf.Bt + wi.z * sf.N;
        rcp = f3{1.0f, 1.0f, 1.0f} / ray_direction;
    }
    return final_color;
}

}  // namespace ptd
