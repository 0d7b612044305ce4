// kernels_mega.hip - the plain method on gfx950: one thread per path sample,
// the whole bounce loop in one kernel (the reference's compute_main,
// src/render/trace.slang:719-1309), plus the query kernels (primary hits,
// closest/any hit of arbitrary rays) and the display kernel.
//
// Launch shape: 256-thread blocks covering 16x16 pixels, each wave an 8x8
// pixel square (the reference's 8x8 workgroup), grid-strided over the tile so
// the global stack-overflow area is sized by the resident grid, not the tile.
#include "kernels.h"

#include <mutex>

namespace ptd {

// Device memory of pt_render_tile_noalloc, for callers that cannot allocate:
// the megakernel's traversal overflow stacks live here instead of in a
// hipMalloc'ed buffer. Each device holds its own copy for the life of the
// process. Host code never passes `scratch` itself (its host-side name is not
// a device address): it takes the per-device address from
// noalloc_scratch_address() below.
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_BYTES / sizeof(uint32_t)];

__device__ __forceinline__ void tile_pixel(uint32_t block, uint32_t tid, uint32_t blocks_x, uint32_t& gx, uint32_t& gy) {
    const uint32_t wave = tid >> 6, lane = tid & 63u;
    gx = (block % blocks_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
    gy = (block / blocks_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
}

// held to 3 waves per SIMD (168 VGPRs): at 2 the whole-path kernel loses a
// third of its throughput
template <bool COUNT>
__global__ __launch_bounds__(256, 3) void k_megakernel(DevScene s, TileArgs a, float4* __restrict__ accum,
                                                     float4* __restrict__ storage, uint32_t* __restrict__ ovf,
                                                     unsigned long long* __restrict__ counters) {
    __shared__ uint32_t lds_stack[MEGA_LDS_DEPTH * 256];
    const uint32_t tid = threadIdx.x;
    Stack st;
    st.lds = as_lds(lds_stack + tid);
    st.lds_stride = 256;
    st.lds_depth = MEGA_LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * 256u;
    st.gtid = blockIdx.x * 256u + tid;
    const uint32_t blocks_x = (a.tile_w + 15u) / 16u, blocks_y = (a.tile_h + 15u) / 16u;
    TraversalCounters ctr = {0, 0};
    uint32_t closest_rays = 0, shadow_rays = 0;
    for (uint32_t blk = blockIdx.x; blk < blocks_x * blocks_y; blk += gridDim.x) {
        uint32_t gx, gy;
        tile_pixel(blk, tid, blocks_x, gx, gy);
        // only the part of the tile inside the image is traced (path_tracer.rs:336-340, trace.slang:729-731)
        if (gx >= a.tile_w || gy >= a.tile_h) continue;
        const uint32_t px = gx + a.origin_x, py = gy + a.origin_y;
        if (px >= a.image_w || py >= a.image_h) continue;
        const size_t idx = (size_t)gy * a.tile_w + gx;
        for (uint32_t k = 0; k < a.samples; ++k) {
            const uint32_t sc = a.first_sample + k;
            f3 L = trace_path<COUNT>(s, px, py, a.image_w, a.image_h, a.max_bounces, sc, st, ctr, closest_rays, shadow_rays);
            // trace.slang:1299-1308
            float4 prev = sc > 1u ? accum[idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            f3 acc = f3{prev.x, prev.y, prev.z} + L;
            accum[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
            if (k + 1u == a.samples && storage) {
                f3 o = PT_EXPOSURE * acc / (float)sc;
                storage[idx] = make_float4(o.x, o.y, o.z, 1.0f);
            }
        }
    }
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)closest_rays);
        atomicAdd(&counters[1], (unsigned long long)shadow_rays);
        atomicAdd(&counters[2], (unsigned long long)ctr.pairs);
        atomicAdd(&counters[3], (unsigned long long)ctr.tris);
    }
}

__global__ __launch_bounds__(256) void k_primary_hits(DevScene s, TileArgs a, int32_t* __restrict__ tri,
                                                      float* __restrict__ tout, uint32_t* __restrict__ ovf) {
    __shared__ uint32_t lds_stack[MEGA_LDS_DEPTH * 256];
    const uint32_t tid = threadIdx.x;
    Stack st;
    st.lds = as_lds(lds_stack + tid);
    st.lds_stride = 256;
    st.lds_depth = MEGA_LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * 256u;
    st.gtid = blockIdx.x * 256u + tid;
    const uint32_t blocks_x = (a.tile_w + 15u) / 16u, blocks_y = (a.tile_h + 15u) / 16u;
    TraversalCounters ctr = {0, 0};
    for (uint32_t blk = blockIdx.x; blk < blocks_x * blocks_y; blk += gridDim.x) {
        uint32_t gx, gy;
        tile_pixel(blk, tid, blocks_x, gx, gy);
        if (gx >= a.tile_w || gy >= a.tile_h) continue;
        const size_t idx = (size_t)gy * a.tile_w + gx;
        const uint32_t px = gx + a.origin_x, py = gy + a.origin_y;
        if (px >= a.image_w || py >= a.image_h) {
            tri[idx] = -1;
            tout[idx] = 0.0f;
            continue;
        }
        Rng rng{ptm::rng_seed(px, py, a.first_sample)};
        f3 o = {s.camera.position[0], s.camera.position[1], s.camera.position[2]};
        f3 d = primary_direction(s, px, py, a.image_w, a.image_h, rng);
        f3 rcp = f3{1.0f, 1.0f, 1.0f} / d;
        float ht, hv, hw;
        tri[idx] = closest_hit<false>(s, o, d, rcp, rng, st, ht, hv, hw, ctr);
        tout[idx] = ht;
    }
}

__global__ __launch_bounds__(256) void k_intersect(DevScene s, const float* __restrict__ rays, uint32_t count,
                                                   int32_t* __restrict__ hit, float* __restrict__ t, float* __restrict__ v,
                                                   float* __restrict__ w, uint32_t* __restrict__ pairs,
                                                   uint32_t* __restrict__ tris, uint32_t* __restrict__ ovf) {
    __shared__ uint32_t lds_stack[MEGA_LDS_DEPTH * 256];
    const uint32_t tid = threadIdx.x;
    Stack st;
    st.lds = as_lds(lds_stack + tid);
    st.lds_stride = 256;
    st.lds_depth = MEGA_LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * 256u;
    st.gtid = blockIdx.x * 256u + tid;
    for (uint32_t i = blockIdx.x * 256u + tid; i < count; i += gridDim.x * 256u) {
        const float* r = rays + (size_t)i * 6;
        f3 o = {r[0], r[1], r[2]}, d = {r[3], r[4], r[5]};
        f3 rcp = f3{1.0f, 1.0f, 1.0f} / d;
        Rng rng{1u};
        TraversalCounters ctr = {0, 0};
        float ht, hv, hw;
        hit[i] = closest_hit<true>(s, o, d, rcp, rng, st, ht, hv, hw, ctr);
        t[i] = ht;
        v[i] = hv;
        w[i] = hw;
        pairs[i] = ctr.pairs;
        tris[i] = ctr.tris;
    }
}

__global__ __launch_bounds__(256) void k_occluded(DevScene s, const float* __restrict__ rays,
                                                  const float* __restrict__ tmax, const int32_t* __restrict__ skip,
                                                  uint32_t count, uint8_t* __restrict__ out, uint32_t* __restrict__ ovf) {
    __shared__ uint32_t lds_stack[MEGA_LDS_DEPTH * 256];
    const uint32_t tid = threadIdx.x;
    Stack st;
    st.lds = as_lds(lds_stack + tid);
    st.lds_stride = 256;
    st.lds_depth = MEGA_LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * 256u;
    st.gtid = blockIdx.x * 256u + tid;
    for (uint32_t i = blockIdx.x * 256u + tid; i < count; i += gridDim.x * 256u) {
        const float* r = rays + (size_t)i * 6;
        Rng rng{1u};
        TraversalCounters ctr = {0, 0};
        out[i] = occluded<false>(s, f3{r[0], r[1], r[2]}, f3{r[3], r[4], r[5]}, tmax[i], skip[i], rng, st, ctr) ? 1 : 0;
    }
}

// display.wgsl:19-46: ACES filmic fit, then the sRGB OETF, into RGBA8
// (round to nearest, as an Rgba8Unorm render target stores it)
__global__ __launch_bounds__(256) void k_display(const float4* __restrict__ storage, uint32_t n, uchar4* __restrict__ out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float4 c = storage[i];
    float in[3] = {c.x, c.y, c.z};
    unsigned char o[3];
    for (int k = 0; k < 3; ++k) {
        float x = in[k];
        float m = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
        m = fminf(fmaxf(m, 0.0f), 1.0f);
        float sv = m > 0.0031308f ? ptm::pow_(m, 1.0f / 2.4f) * 1.055f - 0.055f : m * 12.92f;
        sv = fminf(fmaxf(sv, 0.0f), 1.0f);
        o[k] = (unsigned char)rintf(sv * 255.0f);
    }
    float a = fminf(fmaxf(c.w, 0.0f), 1.0f);
    out[i] = make_uchar4(o[0], o[1], o[2], (unsigned char)rintf(a * 255.0f));
}

}  // namespace ptd

// ------------------------------------------------------------------ launchers
namespace ptcore {

hipError_t noalloc_scratch_address(int device, uint32_t** out) {
    // once per device, with that device current (the caller made it current)
    static std::mutex mu;
    static uint32_t* addr[64] = {};
    *out = nullptr;
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(ptd::scratch));
        if (e != hipSuccess) return e;
        addr[device] = (uint32_t*)p;
    }
    *out = addr[device];
    return hipSuccess;
}

hipError_t launch_megakernel(const ptd::DevScene& s, const ptd::TileArgs& a, float4* accum, float4* storage,
                             uint32_t* ovf, uint32_t grid, unsigned long long* counters, hipStream_t stream) {
    if (counters)
        ptd::k_megakernel<true><<<grid, 256, 0, stream>>>(s, a, accum, storage, ovf, counters);
    else
        ptd::k_megakernel<false><<<grid, 256, 0, stream>>>(s, a, accum, storage, ovf, nullptr);
    return hipGetLastError();
}
hipError_t launch_primary_hits(const ptd::DevScene& s, const ptd::TileArgs& a, int32_t* tri, float* t, uint32_t* ovf,
                               uint32_t grid, hipStream_t stream) {
    ptd::k_primary_hits<<<grid, 256, 0, stream>>>(s, a, tri, t, ovf);
    return hipGetLastError();
}
hipError_t launch_intersect(const ptd::DevScene& s, const float* rays, uint32_t count, int32_t* hit, float* t, float* v,
                            float* w, uint32_t* pairs, uint32_t* tris, uint32_t* ovf, uint32_t grid, hipStream_t stream) {
    ptd::k_intersect<<<grid, 256, 0, stream>>>(s, rays, count, hit, t, v, w, pairs, tris, ovf);
    return hipGetLastError();
}
hipError_t launch_occluded(const ptd::DevScene& s, const float* rays, const float* tmax, const int32_t* skip,
                           uint32_t count, uint8_t* out, uint32_t* ovf, uint32_t grid, hipStream_t stream) {
    ptd::k_occluded<<<grid, 256, 0, stream>>>(s, rays, tmax, skip, count, out, ovf);
    return hipGetLastError();
}
hipError_t launch_display(const float4* storage, uint32_t n, uchar4* out, hipStream_t stream) {
    ptd::k_display<<<(n + 255u) / 256u, 256, 0, stream>>>(storage, n, out);
    return hipGetLastError();
}

}  // namespace ptcore
