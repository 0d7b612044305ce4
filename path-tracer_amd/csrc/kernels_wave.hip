// kernels_wave.hip - wavefront path tracing on gfx950.
//
// Each bounce of the reference's per-thread loop (src/render/trace.slang:757-1297)
// becomes three grid-strided kernels over compacted queues of live paths:
//   extend  closest-hit traversal of the ray queue (LDS short stack)
//   shade   miss/sky, emission with MIS, BSDF setup and E_dielectric, light
//           sample and NEE contribution, BSDF sample, roulette; appends the
//           continuation ray and the shadow ray to their queues
//   shadow  any-hit traversal; adds the NEE contribution when unoccluded
// with `generate` before the first bounce and `accumulate` after the last.
// Queue appends use a wave-level ballot + mbcnt prefix and one atomic per
// wave. Per-path RNG state travels with the path; because opaque traversal
// draws no random numbers (trace.slang:447-451), deferring the shadow ray
// keeps every path's random stream, and the radiance additions keep their
// order (emission, then NEE, then the next bounce), so each path's result
// is bit-identical to the megakernel's. Scenes with alpha-tested materials
// run the shadow test inline in shade to keep the draw order.
#include "wavefront.h"

#include <algorithm>
#include <cstring>

namespace ptd {

constexpr uint32_t WAVE_BLOCK = 256;
constexpr uint32_t MAX_COUNTS = 2 * (PT_MAX_BOUNCES_LIMIT + 2);

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(0xffffffffu, __builtin_amdgcn_mbcnt_lo(0xffffffffu, 0u));
}

// Append for every lane with `pred` set: one atomic per wave, slots by prefix
// count of the ballot. Must be reached by all active lanes of the wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    const uint32_t prefix = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    const uint32_t total = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (total) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane_id() == leader) base = atomicAdd(counter, total);
        base = __shfl(base, leader);
    }
    return base + prefix;
}

// block order: 8x8 pixel squares, row-major over the tile
__device__ __forceinline__ void block_pixel(uint32_t q, uint32_t tile_w, uint32_t& gx, uint32_t& gy) {
    const uint32_t bxc = (tile_w + 7u) >> 3;
    const uint32_t b = q >> 6, l = q & 63u;
    gx = (b % bxc) * 8u + (l & 7u);
    gy = (b / bxc) * 8u + (l >> 3);
}

__device__ __forceinline__ f3 ld3(const float* x, const float* y, const float* z, uint32_t i) { return f3{x[i], y[i], z[i]}; }
__device__ __forceinline__ void st3(float* x, float* y, float* z, uint32_t i, f3 v) {
    x[i] = v.x;
    y[i] = v.y;
    z[i] = v.z;
}

__global__ void k_zero_counts(uint32_t* counts) {
    if (threadIdx.x < MAX_COUNTS) counts[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(WAVE_BLOCK) void k_generate(DevScene s, WaveArgs a, PathState ps, uint32_t* __restrict__ queue,
                                                         uint32_t* __restrict__ counts) {
    const uint32_t n = a.chunk_pixels * a.tile.samples;
    const uint32_t stride = gridDim.x * WAVE_BLOCK;
    for (uint32_t base = blockIdx.x * WAVE_BLOCK; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool active = false;
        if (i < n) {
            const uint32_t k = i / a.chunk_pixels, q = i % a.chunk_pixels;
            uint32_t gx, gy;
            block_pixel(a.chunk_first_pixel + q, a.tile.tile_w, gx, gy);
            const uint32_t px = gx + a.tile.origin_x, py = gy + a.tile.origin_y;
            active = gx < a.tile.tile_w && gy < a.tile.tile_h && px < a.tile.image_w && py < a.tile.image_h;
            ps.lr[i] = 0.0f;
            ps.lg[i] = 0.0f;
            ps.lb[i] = 0.0f;
            if (active) {
                Rng rng{ptm::rng_seed(px, py, a.tile.first_sample + k)};
                f3 d = primary_direction(s, px, py, a.tile.image_w, a.tile.image_h, rng);
                st3(ps.ox, ps.oy, ps.oz, i, f3{s.camera.position[0], s.camera.position[1], s.camera.position[2]});
                st3(ps.dx, ps.dy, ps.dz, i, d);
                st3(ps.tr, ps.tg, ps.tb, i, f3{1.0f, 1.0f, 1.0f});
                ps.prev_pdf[i] = 0.0f;
                ps.rng[i] = rng.s;
            }
        }
        const uint32_t slot = wave_append(&counts[0], active);
        if (active) queue[slot] = i;
    }
}

template <uint32_t LDS_DEPTH, bool COUNT>
__global__ __launch_bounds__(WAVE_BLOCK) void k_extend(DevScene s, PathState ps, const uint32_t* __restrict__ queue,
                                                       const uint32_t* __restrict__ count_ptr, uint32_t* __restrict__ ovf,
                                                       unsigned long long* __restrict__ counters) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * WAVE_BLOCK];
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.lds_stride = WAVE_BLOCK;
    st.lds_depth = LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * WAVE_BLOCK;
    st.gtid = blockIdx.x * WAVE_BLOCK + threadIdx.x;
    const uint32_t count = *count_ptr;
    TraversalCounters ctr = {0, 0};
    for (uint32_t j = blockIdx.x * WAVE_BLOCK + threadIdx.x; j < count; j += gridDim.x * WAVE_BLOCK) {
        const uint32_t p = queue[j];
        const f3 o = ld3(ps.ox, ps.oy, ps.oz, p), d = ld3(ps.dx, ps.dy, ps.dz, p);
        const f3 rcp = f3{1.0f, 1.0f, 1.0f} / d;
        Rng rng{s.all_opaque ? 0u : ps.rng[p]};
        float t, v, w;
        const int32_t h = closest_hit<COUNT>(s, o, d, rcp, rng, st, t, v, w, ctr);
        ps.hit[p] = h;
        ps.ht[p] = t;
        ps.hv[p] = v;
        ps.hw[p] = w;
        if (!s.all_opaque) ps.rng[p] = rng.s;
    }
    if (COUNT) {
        atomicAdd(&counters[2], (unsigned long long)ctr.pairs);
        atomicAdd(&counters[3], (unsigned long long)ctr.tris);
    }
}

template <uint32_t LDS_DEPTH, bool COUNT>
__global__ __launch_bounds__(WAVE_BLOCK) void k_shadow(DevScene s, PathState ps, const uint32_t* __restrict__ queue,
                                                       const uint32_t* __restrict__ count_ptr, uint32_t* __restrict__ ovf,
                                                       unsigned long long* __restrict__ counters) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * WAVE_BLOCK];
    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.lds_stride = WAVE_BLOCK;
    st.lds_depth = LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * WAVE_BLOCK;
    st.gtid = blockIdx.x * WAVE_BLOCK + threadIdx.x;
    const uint32_t count = *count_ptr;
    TraversalCounters ctr = {0, 0};
    for (uint32_t j = blockIdx.x * WAVE_BLOCK + threadIdx.x; j < count; j += gridDim.x * WAVE_BLOCK) {
        const uint32_t p = queue[j];
        const f3 o = ld3(ps.sox, ps.soy, ps.soz, p), d = ld3(ps.sdx, ps.sdy, ps.sdz, p);
        Rng rng{0u};  // opaque scenes only: any-hit draws nothing
        if (!occluded<COUNT>(s, o, d, ps.sdist[p], ps.slight[p], rng, st, ctr)) {
            f3 L = ld3(ps.lr, ps.lg, ps.lb, p) + ld3(ps.scr, ps.scg, ps.scb, p);
            st3(ps.lr, ps.lg, ps.lb, p, L);
        }
    }
    if (COUNT) {
        atomicAdd(&counters[2], (unsigned long long)ctr.pairs);
        atomicAdd(&counters[3], (unsigned long long)ctr.tris);
    }
}

// One bounce of shading for the paths in `queue` (trace.slang:886-1296).
template <bool INLINE_SHADOW>
__global__ __launch_bounds__(WAVE_BLOCK) void k_shade(DevScene s, PathState ps, uint32_t bounce, uint32_t max_bounces,
                                                      const uint32_t* __restrict__ queue, const uint32_t* __restrict__ count_ptr,
                                                      uint32_t* __restrict__ next_queue, uint32_t* __restrict__ next_count,
                                                      uint32_t* __restrict__ shadow_queue, uint32_t* __restrict__ shadow_count,
                                                      uint32_t* __restrict__ ovf) {
    __shared__ uint32_t lds_stack[INLINE_SHADOW ? MEGA_LDS_DEPTH * WAVE_BLOCK : 1];
    Stack st;
    st.lds = lds_stack + (INLINE_SHADOW ? threadIdx.x : 0u);
    st.lds_stride = WAVE_BLOCK;
    st.lds_depth = MEGA_LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * WAVE_BLOCK;
    st.gtid = blockIdx.x * WAVE_BLOCK + threadIdx.x;
    const uint32_t count = *count_ptr;
    const float prob_sky = s.uniforms.light_probabilities[1];
    const uint32_t stride = gridDim.x * WAVE_BLOCK;
    for (uint32_t base = blockIdx.x * WAVE_BLOCK; base < count; base += stride) {
        const uint32_t j = base + threadIdx.x;
        bool cont = false, shadow = false;
        uint32_t p = 0;
        if (j < count) {
            p = queue[j];
            const f3 ray_origin = ld3(ps.ox, ps.oy, ps.oz, p), ray_direction = ld3(ps.dx, ps.dy, ps.dz, p);
            f3 throughput = ld3(ps.tr, ps.tg, ps.tb, p);
            f3 L = ld3(ps.lr, ps.lg, ps.lb, p);
            float previous_bsdf_pdf = ps.prev_pdf[p];
            Rng rng{ps.rng[p]};
            const int32_t hit = ps.hit[p];
            if (hit == -1) {
                float mis = 1.0f;
                if (bounce > 0u) mis = power_heuristic(previous_bsdf_pdf, skybox_light_pdf(s, ray_direction, prob_sky));
                L = L + throughput * evaluate_skybox(s, ray_direction) * mis;
            } else {
                SurfaceHit sf;
                Bsdf b;
                f3 emitted;
                setup_surface(s, (uint32_t)hit, ps.ht[p], ps.hv[p], ps.hw[p], ray_origin, ray_direction, bounce,
                              previous_bsdf_pdf, rng, sf, b, emitted);
                L = L + throughput * emitted;
                const f3 next_origin = sf.hit_point + sf.flat_N * fmaxf(1e-4f, 1e-4f * ptm::length(sf.hit_point));
                LightSample light = sample_light(s, next_origin, rng);
                const f3 lwi = {ptm::dot(sf.Tt, light.direction), ptm::dot(sf.Bt, light.direction), ptm::dot(sf.N, light.direction)};
                if (light.pdf > 0.0f && lwi.z > 0.0f && ptm::any_pos(light.emission)) {
                    f3 f;
                    float pdf;
                    evaluate_lobes(b, sf.wo, lwi, f, pdf);
                    const float weight = bounce < max_bounces ? power_heuristic(light.pdf, pdf) : 1.0f;
                    const f3 contrib = throughput * f * lwi.z * light.emission * (weight / light.pdf);
                    if (INLINE_SHADOW) {
                        TraversalCounters ctr = {0, 0};
                        if (!occluded<false>(s, next_origin, light.direction, light.distance, light.triangle_index, rng, st, ctr))
                            L = L + contrib;
                    } else {
                        shadow = true;
                        st3(ps.sox, ps.soy, ps.soz, p, next_origin);
                        st3(ps.sdx, ps.sdy, ps.sdz, p, light.direction);
                        ps.sdist[p] = light.distance;
                        ps.slight[p] = light.triangle_index;
                        st3(ps.scr, ps.scg, ps.scb, p, contrib);
                    }
                }
                if (bounce < max_bounces) {
                    const f3 wi = sample_bsdf(b, sf.wo, rng);
                    if (wi.z > 0.0f) {
                        f3 f;
                        float pdf;
                        evaluate_lobes(b, sf.wo, wi, f, pdf);
                        if (pdf > 0.0f) {
                            throughput = throughput * (f * wi.z / pdf);
                            cont = true;
                            if (bounce > 3u) {
                                const float p_survive =
                                    ptm::clamp_(fmaxf(throughput.x, fmaxf(throughput.y, throughput.z)), 0.05f, 0.95f);
                                if (rng.next() > p_survive) cont = false;
                                else throughput = throughput / p_survive;
                            }
                            if (cont) {
                                st3(ps.ox, ps.oy, ps.oz, p, next_origin);
                                st3(ps.dx, ps.dy, ps.dz, p, wi.x * sf.Tt + wi.y * sf.Bt + wi.z * sf.N);
                                st3(ps.tr, ps.tg, ps.tb, p, throughput);
                                ps.prev_pdf[p] = pdf;
                                ps.rng[p] = rng.s;
                            }
                        }
                    }
                }
            }
            st3(ps.lr, ps.lg, ps.lb, p, L);
        }
        const uint32_t ns = wave_append(next_count, cont);
        if (cont) next_queue[ns] = p;
        if (!INLINE_SHADOW) {
            const uint32_t ss = wave_append(shadow_count, shadow);
            if (shadow) shadow_queue[ss] = p;
        }
    }
}

// Per pixel, add the chunk's samples in order (trace.slang:1299-1308).
__global__ __launch_bounds__(WAVE_BLOCK) void k_accumulate(WaveArgs a, PathState ps, float4* __restrict__ accum,
                                                           float4* __restrict__ storage) {
    for (uint32_t q = blockIdx.x * WAVE_BLOCK + threadIdx.x; q < a.chunk_pixels; q += gridDim.x * WAVE_BLOCK) {
        uint32_t gx, gy;
        block_pixel(a.chunk_first_pixel + q, a.tile.tile_w, gx, gy);
        if (gx >= a.tile.tile_w || gy >= a.tile.tile_h) continue;
        if (gx + a.tile.origin_x >= a.tile.image_w || gy + a.tile.origin_y >= a.tile.image_h) continue;
        const size_t idx = (size_t)gy * a.tile.tile_w + gx;
        float4 prev = a.tile.first_sample > 1u ? accum[idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        f3 acc = f3{prev.x, prev.y, prev.z};
        for (uint32_t k = 0; k < a.tile.samples; ++k) {
            const uint32_t p = k * a.chunk_pixels + q;
            acc = acc + ld3(ps.lr, ps.lg, ps.lb, p);
        }
        accum[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
        if (storage) {
            f3 o = PT_EXPOSURE * acc / (float)(a.tile.first_sample + a.tile.samples - 1u);
            storage[idx] = make_float4(o.x, o.y, o.z, 1.0f);
        }
    }
}

}  // namespace ptd

// ------------------------------------------------------------------ host orchestration
namespace ptcore {

using namespace ptd;

Wavefront::~Wavefront() {
    if (arena) hipFree(arena);
    if (ovf) hipFree(ovf);
}

pt_status Wavefront::reserve(uint32_t paths) {
    if (paths <= capacity && arena) return PT_OK;
    if (arena) {
        hipStreamSynchronize(ctx->stream);
        hipFree(arena);
        arena = nullptr;
    }
    // 29 per-path 4-byte fields + 3 queues, each 256-byte aligned, + counters
    const size_t field_bytes = ((size_t)paths * 4 + 255) & ~(size_t)255;
    arena_bytes = 32 * field_bytes + ((MAX_COUNTS * 4 + 255) & ~(size_t)255);
    hipError_t e = hipMalloc(&arena, arena_bytes);
    if (e != hipSuccess) {
        capacity = 0;
        return hip_error(e, "wavefront path state");
    }
    capacity = paths;
    char* p = (char*)arena;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 255) & ~(size_t)255;
        return r;
    };
    const size_t f = (size_t)paths * 4;
    float** fields[] = {&ps.ox, &ps.oy, &ps.oz, &ps.dx, &ps.dy, &ps.dz, &ps.tr, &ps.tg, &ps.tb, &ps.lr, &ps.lg, &ps.lb,
                        &ps.prev_pdf, &ps.ht, &ps.hv, &ps.hw, &ps.sox, &ps.soy, &ps.soz, &ps.sdx, &ps.sdy, &ps.sdz,
                        &ps.sdist, &ps.scr, &ps.scg, &ps.scb};
    for (float** fp : fields) *fp = (float*)take(f);
    ps.rng = (uint32_t*)take(f);
    ps.hit = (int32_t*)take(f);
    ps.slight = (int32_t*)take(f);
    for (auto& q : queue) q = (uint32_t*)take(f);
    counts = (uint32_t*)take(MAX_COUNTS * 4);
    if ((size_t)(p - (char*)arena) > arena_bytes) return set_error(PT_ERR_STATE, "path arena overrun");
    return PT_OK;
}

namespace {

template <bool COUNT>
void launch_extend(uint32_t lds_depth, uint32_t grid, hipStream_t st, const DevScene& s, const PathState& ps,
                   const uint32_t* q, const uint32_t* cnt, uint32_t* ovf, unsigned long long* ctr) {
    switch (lds_depth) {
        case 24: k_extend<24, COUNT><<<grid, WAVE_BLOCK, 0, st>>>(s, ps, q, cnt, ovf, ctr); break;
        case 32: k_extend<32, COUNT><<<grid, WAVE_BLOCK, 0, st>>>(s, ps, q, cnt, ovf, ctr); break;
        default: k_extend<16, COUNT><<<grid, WAVE_BLOCK, 0, st>>>(s, ps, q, cnt, ovf, ctr); break;
    }
}
template <bool COUNT>
void launch_shadow(uint32_t lds_depth, uint32_t grid, hipStream_t st, const DevScene& s, const PathState& ps,
                   const uint32_t* q, const uint32_t* cnt, uint32_t* ovf, unsigned long long* ctr) {
    switch (lds_depth) {
        case 24: k_shadow<24, COUNT><<<grid, WAVE_BLOCK, 0, st>>>(s, ps, q, cnt, ovf, ctr); break;
        case 32: k_shadow<32, COUNT><<<grid, WAVE_BLOCK, 0, st>>>(s, ps, q, cnt, ovf, ctr); break;
        default: k_shadow<16, COUNT><<<grid, WAVE_BLOCK, 0, st>>>(s, ps, q, cnt, ovf, ctr); break;
    }
}

struct EventTimer {
    bool on;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    float* sink;
    EventTimer(bool enabled, hipStream_t s, float* out) : on(enabled), st(s), sink(out) {
        if (on) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, st);
        }
    }
    ~EventTimer() {
        if (on) {
            hipEventRecord(b, st);
            hipEventSynchronize(b);
            float ms = 0.0f;
            hipEventElapsedTime(&ms, a, b);
            *sink += ms;
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
    }
};

}  // namespace

pt_status Wavefront::render(Scene& scene, const TileArgs& tile, float4* accum, float4* storage, int profiling,
                            pt_trace_stats* stats) {
    const hipStream_t stream = ctx->stream;
    const uint32_t bw = (tile.tile_w + 7u) / 8u, bh = (tile.tile_h + 7u) / 8u;
    const uint32_t tile_pixels = bw * bh * 64u;
    const uint32_t S = tile.samples;
    uint32_t chunk = std::max<uint32_t>(64u, (max_chunk_paths / S) & ~63u);
    chunk = std::min(chunk, tile_pixels);
    pt_status st = reserve(chunk * S);
    if (st) return st;
    const uint32_t grid = ctx->cu_count * 8u;
    const uint32_t depth = scene.bvh_depth + 1u;
    const uint32_t lds = extend_cfg.lds_depth;
    const size_t need = (size_t)(depth > std::min(lds, MEGA_LDS_DEPTH) ? depth - std::min(lds, MEGA_LDS_DEPTH) : 1u) *
                        grid * WAVE_BLOCK * 4u;
    if (need > ovf_bytes) {
        if (ovf) {
            hipStreamSynchronize(stream);
            hipFree(ovf);
            ovf = nullptr;
        }
        hipError_t e = hipMalloc(&ovf, need);
        if (e != hipSuccess) {
            ovf_bytes = 0;
            return hip_error(e, "wavefront overflow stacks");
        }
        ovf_bytes = need;
    }
    const DevScene& s = scene.view;
    const bool inline_shadow = !s.all_opaque;
    if (profiling) {
        std::memset(stats, 0, sizeof(*stats));
        hipMemsetAsync(ctx->counters, 0, 16 * sizeof(unsigned long long), stream);
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (profiling) {
        hipEventCreate(&t0);
        hipEventCreate(&t1);
        hipEventRecord(t0, stream);
    }
    unsigned long long* ctr = profiling ? ctx->counters : nullptr;
    for (uint32_t first = 0; first < tile_pixels; first += chunk) {
        WaveArgs a;
        a.tile = tile;
        a.chunk_first_pixel = first;
        a.chunk_pixels = std::min(chunk, tile_pixels - first);
        a.tile_pixels = tile_pixels;
        k_zero_counts<<<1, 64, 0, stream>>>(counts);
        const uint32_t gen_grid = std::min<uint32_t>(grid, (a.chunk_pixels * S + WAVE_BLOCK - 1) / WAVE_BLOCK);
        k_generate<<<gen_grid, WAVE_BLOCK, 0, stream>>>(s, a, ps, queue[0], counts);
        for (uint32_t b = 0; b <= tile.max_bounces; ++b) {
            uint32_t* q = queue[b & 1u];
            uint32_t* qn = queue[(b + 1u) & 1u];
            {
                EventTimer tm(profiling, stream, &stats->extend_ms);
                if (profiling) launch_extend<true>(lds, grid, stream, s, ps, q, counts + 2 * b, ovf, ctr);
                else launch_extend<false>(lds, grid, stream, s, ps, q, counts + 2 * b, ovf, ctr);
            }
            {
                EventTimer tm(profiling, stream, &stats->shade_ms);
                if (inline_shadow)
                    k_shade<true><<<grid, WAVE_BLOCK, 0, stream>>>(s, ps, b, tile.max_bounces, q, counts + 2 * b, qn,
                                                                  counts + 2 * (b + 1), queue[2], counts + 2 * b + 1, ovf);
                else
                    k_shade<false><<<grid, WAVE_BLOCK, 0, stream>>>(s, ps, b, tile.max_bounces, q, counts + 2 * b, qn,
                                                                   counts + 2 * (b + 1), queue[2], counts + 2 * b + 1, ovf);
            }
            if (!inline_shadow) {
                EventTimer tm(profiling, stream, &stats->shadow_ms);
                if (profiling) launch_shadow<true>(lds, grid, stream, s, ps, queue[2], counts + 2 * b + 1, ovf, ctr);
                else launch_shadow<false>(lds, grid, stream, s, ps, queue[2], counts + 2 * b + 1, ovf, ctr);
            }
            if (profiling) stats->extend_launches += 1;
        }
        const uint32_t acc_grid = std::min<uint32_t>(grid, (a.chunk_pixels + WAVE_BLOCK - 1) / WAVE_BLOCK);
        k_accumulate<<<acc_grid, WAVE_BLOCK, 0, stream>>>(a, ps, accum, storage);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "wavefront launch");
        if (profiling) {
            uint32_t h[MAX_COUNTS];
            hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, stream);
            hipStreamSynchronize(stream);
            for (uint32_t b = 0; b <= tile.max_bounces; ++b) {
                stats->closest_rays += h[2 * b];
                stats->shadow_rays += h[2 * b + 1];
            }
        }
    }
    if (profiling) {
        hipEventRecord(t1, stream);
        hipEventSynchronize(t1);
        hipEventElapsedTime(&stats->total_ms, t0, t1);
        hipEventDestroy(t0);
        hipEventDestroy(t1);
        unsigned long long h[4];
        hipMemcpy(h, ctx->counters, sizeof(h), hipMemcpyDeviceToHost);
        stats->pair_tests = h[2];
        stats->tri_tests = h[3];
        if (inline_shadow) stats->shadow_rays = 0;  // counted inline, not queued
    }
    return PT_OK;
}

}  // namespace ptcore
