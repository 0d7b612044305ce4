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
//
// Work comes from a job table: one job for a tracer tile, or many jobs of
// different scenes, resolutions and sample counts for a batch. Paths of
// different jobs share the queues; MULTI kernels look up each path's scene.
#include "wavefront.h"

#include <algorithm>
#include <cstring>
#include <new>

namespace ptd {

constexpr uint32_t WAVE_BLOCK = 256;
// Traversal launches: one wave per block, many more waves than a CU holds at
// once (WaveConfig::trace_waves_per_cu; 32 are resident at full occupancy),
// so the hardware dispatcher keeps handing waves to SIMDs whose waves finish
// early instead of leaving them idle behind the slowest ray batch of the grid.
// The shade launches' blocks per CU, their window of queue rounds and the
// lane-split threshold are WaveConfig fields as well (the autotuner sweeps them).
#ifndef PT_TRACE_BLOCK
#define PT_TRACE_BLOCK 64u
#endif
constexpr uint32_t TRACE_BLOCK = PT_TRACE_BLOCK;
constexpr uint32_t MAX_COUNTS = 2 * (PT_MAX_BOUNCES_LIMIT + 2);
constexpr uint32_t SORT_BUCKETS = 256;

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(0xffffffffu, __builtin_amdgcn_mbcnt_lo(0xffffffffu, 0u));
}

// Adds the traversal counters of a wave's lanes to counters[2..5] with one
// atomic per counter per wave (per-lane atomics on four addresses serialise).
__device__ __forceinline__ void add_counters(unsigned long long* counters, const TraversalCounters& cc,
                                             const TraversalCounters& sc) {
    unsigned long long v[4] = {cc.pairs, cc.tris, sc.pairs, sc.tris};
    for (int k = 0; k < 4; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if (lane_id() == 0u)
        for (int k = 0; k < 4; ++k)
            if (v[k]) atomicAdd(&counters[2 + k], v[k]);
}

// Block-level staging of queue appends: lanes append to an LDS buffer
// (wave-aggregated LDS atomic); the block moves the buffer to the global queue
// with one global atomic when it could overflow on the next iteration and at
// the end. This keeps contention on the global counter to a few thousand
// atomics per launch instead of one per wave. All calls are block-uniform.
constexpr uint32_t BQ_CAP = 2048;
struct BlockQueue {
    uint32_t* buf;   // LDS [BQ_CAP]
    uint32_t* n;     // LDS count
    uint32_t* base;  // LDS broadcast of the global slot
};
__device__ __forceinline__ void bq_init(BlockQueue& q) {
    if (threadIdx.x == 0) *q.n = 0u;
    __syncthreads();
}
__device__ __forceinline__ void bq_push(BlockQueue& q, bool pred, uint32_t v) {
    const uint64_t mask = __ballot(pred);
    const uint32_t prefix = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    const uint32_t total = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (total) {
        const int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane_id() == leader) base = atomicAdd(q.n, total);
        base = __shfl(base, leader);
    }
    if (pred) q.buf[base + prefix] = v;
}
// move the staged entries to queue[*count ...] if `force` or the buffer
// could overflow with one more block-sized iteration
__device__ __forceinline__ void bq_flush(BlockQueue& q, uint32_t* __restrict__ count, uint32_t* __restrict__ queue,
                                         bool force) {
    __syncthreads();
    const uint32_t cnt = *q.n;
    if (cnt != 0u && (force || cnt > BQ_CAP - WAVE_BLOCK)) {
        if (threadIdx.x == 0) *q.base = atomicAdd(count, cnt);
        __syncthreads();
        const uint32_t b = *q.base;
        for (uint32_t i = threadIdx.x; i < cnt; i += WAVE_BLOCK) queue[b + i] = q.buf[i];
        __syncthreads();
        if (threadIdx.x == 0) *q.n = 0u;
        __syncthreads();
    }
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

// scene and job of path p: the kernel arguments for a single job, the tables otherwise
template <bool MULTI>
__device__ __forceinline__ const DevScene& scene_of(const DevScene& s0, const DevScene* scenes, const WaveJob* jobs,
                                                    uint32_t job) {
    if (MULTI) return scenes[jobs[job].scene];
    return s0;
}
template <bool MULTI>
__device__ __forceinline__ const WaveJob& job_of(const WaveJob& j0, const WaveJob* jobs, uint32_t job) {
    if (MULTI) return jobs[job];
    return j0;
}

__global__ void k_zero_counts(uint32_t* counts) {
    if (threadIdx.x < MAX_COUNTS) counts[threadIdx.x] = 0u;
}

// unit of path i: the last unit whose offset is <= i (a division when all
// units have the same size, else a binary search over the offsets)
__device__ __forceinline__ uint32_t find_unit(const ChunkUnits& u, uint32_t i) {
    if (u.unit_paths) return i / u.unit_paths;
    uint32_t lo = 0, hi = u.n_units;  // offset[lo] <= i < offset[hi]
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (u.offset[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <bool MULTI>
__global__ __launch_bounds__(WAVE_BLOCK) void k_generate(DevScene s0, WaveJob j0, const DevScene* __restrict__ scenes,
                                                         const WaveJob* __restrict__ jobs, ChunkUnits units, PathState ps,
                                                         uint32_t* __restrict__ queue, uint32_t* __restrict__ counts) {
    __shared__ uint32_t q_buf[BQ_CAP];
    __shared__ uint32_t q_n, q_base;
    const uint32_t n = units.n_paths;
    const uint32_t stride = gridDim.x * WAVE_BLOCK;
    if (units.all_active) {
        // every pixel of every unit is in its tile and image: the queue is
        // 0 .. n-1 in order (what the appends below produce when no path is
        // inactive, in a scattered order). For one job's consecutive blocks the
        // unit of path i is i / unit_paths and no table is read.
        for (uint32_t i = blockIdx.x * WAVE_BLOCK + threadIdx.x; i < n; i += stride) {
            uint32_t ji, block, rel;
            if (units.contiguous) {
                const uint32_t u = i / units.unit_paths;
                ji = units.job0;
                block = units.block0 + u;
                rel = i - u * units.unit_paths;
            } else {
                const uint32_t u = find_unit(units, i);
                ji = units.job[u];
                block = units.block[u];
                rel = i - units.offset[u];
            }
            const WaveJob& jb = job_of<MULTI>(j0, jobs, ji);
            const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, ji);
            const uint32_t k = rel >> 6, lane = rel & 63u;
            uint32_t gx, gy;
            block_pixel(block * 64u + lane, jb.tile_w, gx, gy);
            const uint32_t px = gx + jb.origin_x, py = gy + jb.origin_y;
            Rng rng{ptm::rng_seed(px, py, jb.first_sample + k)};
            f3 d = primary_direction(s, px, py, jb.image_w, jb.image_h, rng);
            st3(ps.dx, ps.dy, ps.dz, i, d);
            ps.rng[i] = rng.s;
            if (MULTI) ps.job[i] = ji;
            queue[i] = i;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) counts[0] = n;
        return;
    }
    BlockQueue bq{q_buf, &q_n, &q_base};
    bq_init(bq);
    for (uint32_t base = blockIdx.x * WAVE_BLOCK; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool active = false;
        if (i < n) {
            const uint32_t u = find_unit(units, i);
            const uint32_t ji = units.job[u];
            const WaveJob& jb = job_of<MULTI>(j0, jobs, ji);
            const uint32_t rel = i - units.offset[u];
            const uint32_t k = rel >> 6, lane = rel & 63u;
            uint32_t gx, gy;
            block_pixel(units.block[u] * 64u + lane, jb.tile_w, gx, gy);
            const uint32_t px = gx + jb.origin_x, py = gy + jb.origin_y;
            active = gx < jb.tile_w && gy < jb.tile_h && px < jb.image_w && py < jb.image_h;
            // a camera ray: its origin, the unit throughput, the zero radiance
            // and the unused previous pdf are implied at bounce 0, not stored
            if (active) {
                const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, ji);
                Rng rng{ptm::rng_seed(px, py, jb.first_sample + k)};
                f3 d = primary_direction(s, px, py, jb.image_w, jb.image_h, rng);
                st3(ps.dx, ps.dy, ps.dz, i, d);
                ps.rng[i] = rng.s;
                if (MULTI) ps.job[i] = ji;
            }
        }
        bq_push(bq, active, i);
        bq_flush(bq, &counts[0], queue, false);
    }
    bq_flush(bq, &counts[0], queue, true);
}

// Traversal of one bounce: entries [0, n_closest) of the launch are the
// closest-hit rays of `closest_queue`, entries [n_closest, n_closest +
// n_shadow) the shadow rays of the previous bounce's `shadow_queue`. Merging
// the two keeps the launches large; a shadow ray's NEE term is added to the
// path's radiance here, after that bounce's emission (shade b) and before the
// next emission (shade b + 1), which is the shader's order.
// minimum waves per SIMD the traversal kernel is compiled for. At 7 the
// multi-scene instance fits 72 VGPRs instead of 79 and runs 7 waves instead
// of 6 (the 64-scene batch: 2.317 -> 2.353 G path samples/s); the
// single-scene instance, whose scalar registers cap it at 7 waves, is
// unchanged in speed.
#ifndef PT_TRACE_WAVES_PER_EU
#define PT_TRACE_WAVES_PER_EU 7
#endif
template <uint32_t LDS_DEPTH, bool COUNT, bool MULTI>
__global__ __launch_bounds__(TRACE_BLOCK, PT_TRACE_WAVES_PER_EU) void k_trace(DevScene s0, const DevScene* __restrict__ scenes,
                                                      const WaveJob* __restrict__ jobs, PathState ps,
                                                      const uint32_t* __restrict__ closest_queue,
                                                      const uint32_t* __restrict__ closest_count, uint32_t first,
                                                      const uint32_t* __restrict__ shadow_queue,
                                                      const uint32_t* __restrict__ shadow_count,
                                                      uint32_t* __restrict__ ovf, unsigned long long* __restrict__ counters) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * TRACE_BLOCK];
    Stack st;
    st.lds = as_lds(lds_stack + threadIdx.x);
    st.lds_stride = TRACE_BLOCK;
    st.lds_depth = LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * TRACE_BLOCK;
    st.gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    const uint32_t nc = closest_queue ? *closest_count : 0u;
    const uint32_t ns = shadow_queue ? *shadow_count : 0u;
    TraversalCounters cc = {0, 0}, sc = {0, 0};
    for (uint32_t j = blockIdx.x * TRACE_BLOCK + threadIdx.x; j < nc + ns; j += gridDim.x * TRACE_BLOCK) {
        if (j < nc) {
            const uint32_t p = closest_queue[j];
            const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
            const f3 o = first ? camera_origin(s) : ld3(ps.ox, ps.oy, ps.oz, p), d = ld3(ps.dx, ps.dy, ps.dz, p);
            const f3 rcp = f3{1.0f, 1.0f, 1.0f} / d;
            const bool opaque = s.all_opaque != 0u;
            Rng rng{opaque ? 0u : ps.rng[p]};
            float t, v, w;
            const int32_t h = closest_hit<COUNT>(s, o, d, rcp, rng, st, t, v, w, cc);
            ps.hit[p] = h;
            ps.ht[p] = t;
            ps.hv[p] = v;
            ps.hw[p] = w;
            if (!opaque) ps.rng[p] = rng.s;
        } else {
            const uint32_t p = shadow_queue[j - nc];
            const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
            const f3 o = ld3(ps.sox, ps.soy, ps.soz, p), d = ld3(ps.sdx, ps.sdy, ps.sdz, p);
            Rng rng{0u};  // only opaque scenes queue shadow rays: any-hit draws nothing
            if (!occluded<COUNT>(s, o, d, ps.sdist[p], ps.slight[p], rng, st, sc)) {
                f3 L = ld3(ps.lr, ps.lg, ps.lb, p) + ld3(ps.scr, ps.scg, ps.scb, p);
                st3(ps.lr, ps.lg, ps.lb, p, L);
            }
        }
    }
    if (COUNT) {
        add_counters(counters, cc, sc);
    }
}

// The same work as k_trace with lane refill: each lane walks its own ray one
// step per iteration and, when it finishes, stores the result and starts its
// next ray (grid-stride order) in the same iteration, so a wave runs for about
// the average ray cost of its lanes instead of its slowest ray per batch.
template <uint32_t LDS_DEPTH, bool COUNT, bool MULTI>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_refill(DevScene s0, const DevScene* __restrict__ scenes,
                                                             const WaveJob* __restrict__ jobs, PathState ps,
                                                             const uint32_t* __restrict__ closest_queue,
                                                             const uint32_t* __restrict__ closest_count, uint32_t first,
                                                             const uint32_t* __restrict__ shadow_queue,
                                                             const uint32_t* __restrict__ shadow_count,
                                                             uint32_t* __restrict__ ovf,
                                                             unsigned long long* __restrict__ counters) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * TRACE_BLOCK];
    Stack st;
    st.lds = as_lds(lds_stack + threadIdx.x);
    st.lds_stride = TRACE_BLOCK;
    st.lds_depth = LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * TRACE_BLOCK;
    st.gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    const uint32_t nc = closest_queue ? *closest_count : 0u;
    const uint32_t ns = shadow_queue ? *shadow_count : 0u;
    const uint32_t total = nc + ns;
    const uint32_t stride = gridDim.x * TRACE_BLOCK;
    TraversalCounters cc = {0, 0}, sc = {0, 0};
    uint32_t j = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    RayState r;
    uint32_t p = 0;
    Rng rng{0u};
    bool opaque = true;
    auto begin = [&](uint32_t jj) {
        if (jj < nc) {
            p = closest_queue[jj];
            const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
            opaque = s.all_opaque != 0u;
            rng.s = opaque ? 0u : ps.rng[p];
            ray_begin(s, r, first ? camera_origin(s) : ld3(ps.ox, ps.oy, ps.oz, p), ld3(ps.dx, ps.dy, ps.dz, p), false,
                      0.0f, -1, st);
        } else {
            p = shadow_queue[jj - nc];
            const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
            opaque = true;  // only opaque scenes queue shadow rays
            rng.s = 0u;
            ray_begin(s, r, ld3(ps.sox, ps.soy, ps.soz, p), ld3(ps.sdx, ps.sdy, ps.sdz, p), true, ps.sdist[p],
                      ps.slight[p], st);
        }
    };
    bool have = j < total;
    if (have) begin(j);
    while (have) {
        const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
        const bool done = r.shadow ? ray_step<COUNT>(s, r, rng, st, sc) : ray_step<COUNT>(s, r, rng, st, cc);
        if (done) {
            if (r.shadow) {
                if (r.hit == 0) {
                    f3 L = ld3(ps.lr, ps.lg, ps.lb, p) + ld3(ps.scr, ps.scg, ps.scb, p);
                    st3(ps.lr, ps.lg, ps.lb, p, L);
                }
            } else {
                ps.hit[p] = r.hit;
                ps.ht[p] = r.hit_t;
                ps.hv[p] = r.hit_v;
                ps.hw[p] = r.hit_w;
                if (!opaque) ps.rng[p] = rng.s;
            }
            j += stride;
            have = j < total;
            if (have) begin(j);
        }
    }
    if (COUNT) {
        add_counters(counters, cc, sc);
    }
}

// The same work as k_trace, one phase for the closest rays and one for the
// shadow rays, with threshold refill: a lane whose ray is finished idles until
// fewer than PT_REFILL_MIN_ACTIVE lanes of its wave are still traversing; then
// every idle lane starts its next ray (grid-stride order) together, so the
// setup of new rays runs once for many lanes and waves stay mostly full.
#ifndef PT_REFILL_MIN_ACTIVE
#define PT_REFILL_MIN_ACTIVE 32
#endif
template <bool SHADOW, bool COUNT, bool MULTI>
__device__ __forceinline__ void trace_phase(const DevScene& s0, const DevScene* __restrict__ scenes,
                                            const WaveJob* __restrict__ jobs, const PathState& ps,
                                            const uint32_t* __restrict__ queue, uint32_t n, bool first,
                                            Stack& st, TraversalCounters& ctr) {
    const uint32_t stride = gridDim.x * TRACE_BLOCK;
    uint32_t j = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    bool have = false, opaque = true;
    uint32_t p = 0;
    RayState r;
    Rng rng{0u};
    for (;;) {
        if (__popcll(__ballot(have)) < PT_REFILL_MIN_ACTIVE) {
            if (!have && j < n) {
                p = queue[j];
                j += stride;
                have = true;
                const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
                if (SHADOW) {
                    // only opaque scenes queue shadow rays: any-hit draws nothing
                    ray_begin(s, r, ld3(ps.sox, ps.soy, ps.soz, p), ld3(ps.sdx, ps.sdy, ps.sdz, p), true, ps.sdist[p],
                              ps.slight[p], st);
                } else {
                    opaque = s.all_opaque != 0u;
                    rng.s = opaque ? 0u : ps.rng[p];
                    ray_begin(s, r, first ? camera_origin(s) : ld3(ps.ox, ps.oy, ps.oz, p), ld3(ps.dx, ps.dy, ps.dz, p),
                              false, 0.0f, -1, st);
                }
            }
            if (__ballot(have) == 0ull) break;
        }
        if (have) {
            const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
            if (ray_advance<SHADOW, COUNT>(s, r, rng, st, ctr)) {
                have = false;
                if (SHADOW) {
                    if (r.hit == 0) {
                        f3 L = ld3(ps.lr, ps.lg, ps.lb, p) + ld3(ps.scr, ps.scg, ps.scb, p);
                        st3(ps.lr, ps.lg, ps.lb, p, L);
                    }
                } else {
                    ps.hit[p] = r.hit;
                    ps.ht[p] = r.hit_t;
                    ps.hv[p] = r.hit_v;
                    ps.hw[p] = r.hit_w;
                    if (!opaque) ps.rng[p] = rng.s;
                }
            }
        }
    }
}

template <uint32_t LDS_DEPTH, bool COUNT, bool MULTI>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_fetch(DevScene s0, const DevScene* __restrict__ scenes,
                                                            const WaveJob* __restrict__ jobs, PathState ps,
                                                            const uint32_t* __restrict__ closest_queue,
                                                            const uint32_t* __restrict__ closest_count, uint32_t first,
                                                            const uint32_t* __restrict__ shadow_queue,
                                                            const uint32_t* __restrict__ shadow_count,
                                                            uint32_t* __restrict__ ovf,
                                                            unsigned long long* __restrict__ counters) {
    __shared__ uint32_t lds_stack[LDS_DEPTH * TRACE_BLOCK];
    Stack st;
    st.lds = as_lds(lds_stack + threadIdx.x);
    st.lds_stride = TRACE_BLOCK;
    st.lds_depth = LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * TRACE_BLOCK;
    st.gtid = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    TraversalCounters cc = {0, 0}, sc = {0, 0};
    if (closest_queue)
        trace_phase<false, COUNT, MULTI>(s0, scenes, jobs, ps, closest_queue, *closest_count, first != 0u, st, cc);
    if (shadow_queue)
        trace_phase<true, COUNT, MULTI>(s0, scenes, jobs, ps, shadow_queue, *shadow_count, false, st, sc);
    if (COUNT) {
        add_counters(counters, cc, sc);
    }
}

// One bounce of shading for the paths in `queue` (trace.slang:886-1296).
// Paths of scenes with alpha-tested materials test their shadow ray inline
// (the any-hit draws random numbers between NEE and BSDF sampling); only the
// INLINE_SHADOW instance carries that traversal and its LDS stack, so launches
// over opaque scenes keep the smaller register and LDS footprint.
// minimum waves per SIMD the shade kernel is compiled for (its register
// budget): 3 fits every instance in 168 VGPRs. The opaque instances need no
// more; the multi-scene inline-shadow ones would otherwise take 169-181 and
// run at 2 waves (a batch holding an alpha-tested scene: 1.25 -> 1.36 G path
// samples/s at 3, with a few bytes of scratch).
#ifndef PT_SHADE_WAVES_PER_EU
#define PT_SHADE_WAVES_PER_EU 3
#endif
template <bool MULTI, bool INLINE_SHADOW, uint32_t WIN>
__global__ __launch_bounds__(WAVE_BLOCK, PT_SHADE_WAVES_PER_EU) void k_shade(DevScene s0, const DevScene* __restrict__ scenes,
                                                      const WaveJob* __restrict__ jobs, WaveJob j0, PathState ps,
                                                      uint32_t bounce, const uint32_t* __restrict__ queue,
                                                      const uint32_t* __restrict__ count_ptr,
                                                      uint32_t* __restrict__ next_queue, uint32_t* __restrict__ next_count,
                                                      uint32_t* __restrict__ shadow_queue, uint32_t* __restrict__ shadow_count,
                                                      uint32_t* __restrict__ ovf) {
    __shared__ uint32_t lds_stack[INLINE_SHADOW ? MEGA_LDS_DEPTH * WAVE_BLOCK : 1];
    Stack st;
    st.lds = as_lds(lds_stack + (INLINE_SHADOW ? threadIdx.x : 0u));
    st.lds_stride = WAVE_BLOCK;
    st.lds_depth = MEGA_LDS_DEPTH;
    st.ovf = ovf;
    st.ovf_stride = gridDim.x * WAVE_BLOCK;
    st.gtid = blockIdx.x * WAVE_BLOCK + threadIdx.x;
    __shared__ uint32_t nq_buf[BQ_CAP], sq_buf[BQ_CAP];
    __shared__ uint32_t nq_n, nq_base, sq_n, sq_base;
    BlockQueue nq{nq_buf, &nq_n, &nq_base}, sq{sq_buf, &sq_n, &sq_base};
    bq_init(nq);
    bq_init(sq);
    // The block takes WIN rounds of WAVE_BLOCK queue entries at a time
    // and reorders them in LDS: paths whose ray hit a surface first, then the
    // missed ones, each in queue order. A wave then runs either the surface
    // work (BSDF setup, E_dielectric, NEE, sampling: thousands of
    // instructions) or the sky lookup, instead of idling its missed lanes
    // through the surface work. Each path's computation is unchanged.
    constexpr uint32_t NW = WAVE_BLOCK / 64u;
    __shared__ uint32_t win[WIN * WAVE_BLOCK];
    __shared__ uint32_t win_hits[WIN * NW], win_misses[WIN * NW];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t count = *count_ptr;
    // each block takes a contiguous run of windows, so what it appends comes
    // from one stretch of the queue: with blocks finishing roughly in launch
    // order, the next queues keep the order of this one, and neighbouring
    // waves of the next traversal trace neighbouring pixels
    const uint32_t span = WAVE_BLOCK * WIN;
    const uint32_t windows = (count + span - 1) / span;
    const uint32_t per_block = (windows + gridDim.x - 1) / gridDim.x;
    const uint32_t w_begin = blockIdx.x * per_block, w_end = min(w_begin + per_block, windows);
    for (uint32_t wi = w_begin; wi < w_end; ++wi) {
        const uint32_t wbase = wi * span;
        uint32_t pk[WIN];
        bool hk[WIN], mk[WIN];
#pragma unroll
        for (uint32_t k = 0; k < WIN; ++k) {
            const uint32_t j = wbase + k * WAVE_BLOCK + threadIdx.x;
            const bool valid = j < count;
            pk[k] = valid ? queue[j] : 0u;
            const int32_t h = valid ? ps.hit[pk[k]] : 0;
            hk[k] = valid && h != -1;
            mk[k] = valid && h == -1;
            const uint32_t nh = (uint32_t)__popcll(__ballot(hk[k])), nm = (uint32_t)__popcll(__ballot(mk[k]));
            if (lane_id() == 0u) {
                win_hits[k * NW + wave] = nh;
                win_misses[k * NW + wave] = nm;
            }
        }
        __syncthreads();
        uint32_t off_h[WIN], off_m[WIN], run_h = 0, run_m = 0;
#pragma unroll
        for (uint32_t i = 0; i < WIN * NW; ++i) {
            if (i % NW == wave) {
                off_h[i / NW] = run_h;
                off_m[i / NW] = run_m;
            }
            run_h += win_hits[i];
            run_m += win_misses[i];
        }
        const uint32_t n = run_h + run_m;
#pragma unroll
        for (uint32_t k = 0; k < WIN; ++k) {
            const uint64_t bh = __ballot(hk[k]), bm = __ballot(mk[k]);
            if (hk[k]) win[off_h[k] + __builtin_amdgcn_mbcnt_hi((uint32_t)(bh >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bh, 0u))] = pk[k];
            if (mk[k])
                win[run_h + off_m[k] + __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u))] = pk[k];
        }
        __syncthreads();
        for (uint32_t k = 0; k < WIN; ++k) {
            const uint32_t idx = k * WAVE_BLOCK + threadIdx.x;
            bool cont = false, shadow = false;
            uint32_t p = 0;
            if (idx < n) {
                p = win[idx];
                const uint32_t ji = MULTI ? ps.job[p] : 0u;
                const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, ji);
                const uint32_t max_bounces = job_of<MULTI>(j0, jobs, ji).max_bounces;
                const bool first = bounce == 0u;  // camera ray: state implied (see k_generate)
                const f3 ray_origin = first ? camera_origin(s) : ld3(ps.ox, ps.oy, ps.oz, p);
                const f3 ray_direction = ld3(ps.dx, ps.dy, ps.dz, p);
                f3 throughput = first ? f3{1.0f, 1.0f, 1.0f} : ld3(ps.tr, ps.tg, ps.tb, p);
                f3 L = first ? f3{0.0f, 0.0f, 0.0f} : ld3(ps.lr, ps.lg, ps.lb, p);
                const float previous_bsdf_pdf = first ? 0.0f : ps.prev_pdf[p];
                Rng rng{ps.rng[p]};
                const int32_t hit = ps.hit[p];
                if (hit == -1) {
                    float mis = 1.0f;
                    if (bounce > 0u)
                        mis = power_heuristic(previous_bsdf_pdf, skybox_light_pdf(s, ray_direction, s.uniforms.light_probabilities[1]));
                    L = L + throughput * evaluate_skybox(s, ray_direction) * mis;
                } else {
                    SurfaceHit sf;
                    Bsdf b;
                    f3 emission;
                    float emission_mis;
                    setup_surface(s, (uint32_t)hit, ps.ht[p], ps.hv[p], ps.hw[p], ray_origin, ray_direction, bounce,
                                  previous_bsdf_pdf, rng, sf, b, emission, emission_mis);
                    L = L + throughput * emission * emission_mis;
                    const f3 next_origin = sf.hit_point + sf.flat_N * fmaxf(1e-4f, 1e-4f * ptm::length(sf.hit_point));
                    LightSample light = sample_light(s, next_origin, rng);
                    const f3 lwi = {ptm::dot(sf.Tt, light.direction), ptm::dot(sf.Bt, light.direction), ptm::dot(sf.N, light.direction)};
                    if (light.pdf > 0.0f && lwi.z > 0.0f && ptm::any_pos(light.emission)) {
                        f3 f;
                        float pdf;
                        evaluate_lobes(b, sf.wo, lwi, f, pdf);
                        const float weight = bounce < max_bounces ? power_heuristic(light.pdf, pdf) : 1.0f;
                        const f3 contrib = throughput * f * lwi.z * light.emission * (weight / light.pdf);
                        if (INLINE_SHADOW && !s.all_opaque) {
                            TraversalCounters ctr = {0, 0};
                            if (!occluded<false, false>(s, next_origin, light.direction, light.distance, light.triangle_index, rng, st, ctr))
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
            bq_push(nq, cont, p);
            bq_push(sq, shadow, p);
            bq_flush(nq, next_count, next_queue, false);
            bq_flush(sq, shadow_count, shadow_queue, false);
        }
    }
    bq_flush(nq, next_count, next_queue, true);
    bq_flush(sq, shadow_count, shadow_queue, true);
}

// Per pixel of each unit, add the job's samples in order (trace.slang:1299-1308).
template <bool MULTI>
__global__ __launch_bounds__(WAVE_BLOCK) void k_accumulate(WaveJob j0, const WaveJob* __restrict__ jobs, ChunkUnits units,
                                                           PathState ps) {
    const uint32_t n = units.n_units * 64u;
    for (uint32_t t = blockIdx.x * WAVE_BLOCK + threadIdx.x; t < n; t += gridDim.x * WAVE_BLOCK) {
        const uint32_t u = t >> 6, lane = t & 63u;
        const WaveJob& jb = job_of<MULTI>(j0, jobs, units.job[u]);
        uint32_t gx, gy;
        block_pixel(units.block[u] * 64u + lane, jb.tile_w, gx, gy);
        if (gx >= jb.tile_w || gy >= jb.tile_h) continue;
        if (gx + jb.origin_x >= jb.image_w || gy + jb.origin_y >= jb.image_h) continue;
        const size_t idx = (size_t)gy * jb.tile_w + gx;
        float4 prev = jb.first_sample > 1u ? jb.accum[idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        f3 acc = f3{prev.x, prev.y, prev.z};
        const uint32_t base = units.offset[u] + lane;
        // unrolled so that several samples' loads are in flight; the sum stays in sample order
#pragma unroll 8
        for (uint32_t k = 0; k < jb.samples; ++k) acc = acc + ld3(ps.lr, ps.lg, ps.lb, base + k * 64u);
        jb.accum[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
        if (jb.storage) {
            f3 o = PT_EXPOSURE * acc / (float)(jb.first_sample + jb.samples - 1u);
            jb.storage[idx] = make_float4(o.x, o.y, o.z, 1.0f);
        }
    }
}

// ------------------------------------------------------------------ queue sorting
// Block-local counting sort of a queue by a small key (direction octant with
// the scene in the high bits, or hit material): each block sorts contiguous
// segments of SORT_SEG entries in LDS, so the lanes of a wave see coherent
// rays / materials while each segment stays where it was in the queue (the
// path state it points at keeps its locality).
constexpr uint32_t SORT_SEG = 2048;

template <bool MULTI>
__device__ __forceinline__ uint32_t sort_key(uint32_t key_kind, const DevScene& s0, const DevScene* scenes,
                                             const WaveJob* jobs, const PathState& ps, uint32_t p) {
    if (key_kind == SORT_OCTANT) {
        uint32_t oct = (ps.dx[p] < 0.0f ? 1u : 0u) | (ps.dy[p] < 0.0f ? 2u : 0u) | (ps.dz[p] < 0.0f ? 4u : 0u);
        uint32_t sc = MULTI ? (jobs[ps.job[p]].scene & 31u) : 0u;
        return (sc << 3) | oct;
    }
    const int32_t h = ps.hit[p];
    if (h < 0) return SORT_BUCKETS - 1u;
    const DevScene& s = scene_of<MULTI>(s0, scenes, jobs, MULTI ? ps.job[p] : 0u);
    return min(s.attr[h].material, SORT_BUCKETS - 2u);
}

template <bool MULTI>
__global__ __launch_bounds__(WAVE_BLOCK) void k_sort_local(uint32_t key_kind, DevScene s0, const DevScene* __restrict__ scenes,
                                                           const WaveJob* __restrict__ jobs, PathState ps,
                                                           const uint32_t* __restrict__ queue,
                                                           const uint32_t* __restrict__ count_ptr, uint32_t* __restrict__ out) {
    __shared__ uint32_t vals[SORT_SEG];
    __shared__ uint8_t keys[SORT_SEG];
    __shared__ uint32_t hist[SORT_BUCKETS];
    const uint32_t count = *count_ptr;
    const uint32_t t = threadIdx.x;
    for (uint32_t seg = blockIdx.x * SORT_SEG; seg < count; seg += gridDim.x * SORT_SEG) {
        const uint32_t n = min(SORT_SEG, count - seg);
        hist[t] = 0u;
        __syncthreads();
        for (uint32_t i = t; i < n; i += WAVE_BLOCK) {
            const uint32_t p = queue[seg + i];
            const uint32_t k = sort_key<MULTI>(key_kind, s0, scenes, jobs, ps, p);
            vals[i] = p;
            keys[i] = (uint8_t)k;
            atomicAdd(&hist[k], 1u);
        }
        __syncthreads();
        // exclusive scan of the 256 bucket counts (one entry per thread)
        const uint32_t mine = hist[t];
        for (uint32_t off = 1; off < SORT_BUCKETS; off <<= 1) {
            const uint32_t add = t >= off ? hist[t - off] : 0u;
            __syncthreads();
            hist[t] += add;
            __syncthreads();
        }
        hist[t] -= mine;
        __syncthreads();
        for (uint32_t i = t; i < n; i += WAVE_BLOCK) out[seg + atomicAdd(&hist[keys[i]], 1u)] = vals[i];
        __syncthreads();
    }
}

}  // namespace ptd

// ------------------------------------------------------------------ host orchestration
namespace ptcore {

using namespace ptd;

namespace {

// Records a start/stop event pair around a kernel on the stream it runs on;
// no synchronisation until the log is read.
struct EventScope {
    EventLog* log;
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    EventScope(EventLog* l, int k, hipStream_t s) : log(l && l->records(k) ? l : nullptr), kind(k), st(s) {
        if (log) {
            log->take(a, b);
            if (a && hipEventRecord(a, st) != hipSuccess) a = nullptr;
        }
    }
    ~EventScope() {
        if (log) {
            if (b && hipEventRecord(b, st) != hipSuccess) b = nullptr;
            log->pairs.push_back({kind, a, b});
        }
    }
};

template <bool COUNT, bool MULTI>
void launch_trace(uint32_t lds, uint32_t refill, uint32_t grid, hipStream_t st, const DevScene& s0,
                  const DevScene* scenes, const WaveJob* jobs, const PathState& ps, const uint32_t* cq,
                  const uint32_t* cc, uint32_t first, const uint32_t* sq, const uint32_t* scnt, uint32_t* ovf,
                  unsigned long long* ctr) {
#define PT_TRACE_LAUNCH(K, D) \
    K<D, COUNT, MULTI><<<grid, TRACE_BLOCK, 0, st>>>(s0, scenes, jobs, ps, cq, cc, first, sq, scnt, ovf, ctr)
    if (refill == TRAVERSAL_REFILL) {
        switch (lds) {
            case 24: PT_TRACE_LAUNCH(k_trace_refill, 24); break;
            case 32: PT_TRACE_LAUNCH(k_trace_refill, 32); break;
            default: PT_TRACE_LAUNCH(k_trace_refill, 16); break;
        }
    } else if (refill == TRAVERSAL_FETCH) {
        switch (lds) {
            case 24: PT_TRACE_LAUNCH(k_trace_fetch, 24); break;
            case 32: PT_TRACE_LAUNCH(k_trace_fetch, 32); break;
            default: PT_TRACE_LAUNCH(k_trace_fetch, 16); break;
        }
    } else {
        switch (lds) {
            case 24: PT_TRACE_LAUNCH(k_trace, 24); break;
            case 32: PT_TRACE_LAUNCH(k_trace, 32); break;
            default: PT_TRACE_LAUNCH(k_trace, 16); break;
        }
    }
#undef PT_TRACE_LAUNCH
}

struct Launch {
    uint32_t grid;
    uint32_t trace_grid;  // blocks of the traversal launches
    uint32_t shade_grid;  // blocks of the shade launches
    hipStream_t stream;
    const DevScene* s0;
    const DevScene* scenes;
    const WaveJob* jobs;
    const WaveJob* j0;
    ChunkUnits units;
    bool any_alpha;  // some scene has alpha-tested materials: shade tests those shadow rays inline
};

template <bool MULTI>
void sort_queue(Wavefront& w, const Launch& L, uint32_t key, const uint32_t* in, const uint32_t* cnt, uint32_t* out) {
    k_sort_local<MULTI><<<L.grid, WAVE_BLOCK, 0, L.stream>>>(key, *L.s0, L.scenes, L.jobs, w.ps, in, cnt, out);
}

// One chunk: generate; per bounce b: trace (closest rays of b + shadow rays
// of b - 1), shade b; a last shadow-only trace; accumulate. Ray queues rotate
// among queue[0], queue[1], queue[3] (cur, next, sort scratch); shadow rays
// use queue[2], consumed by the trace that follows the shade producing them.
template <bool MULTI>
void run_chunk(Wavefront& w, const Launch& L, uint32_t max_bounces, bool counting, EventLog* lg,
               unsigned long long* ctr) {
    const hipStream_t stream = L.stream;
    const uint32_t grid = L.grid;
    const PathState& ps = w.ps;
    k_zero_counts<<<1, 64, 0, stream>>>(w.counts);
    const uint32_t gen_grid = std::min<uint32_t>(grid, (L.units.n_paths + WAVE_BLOCK - 1) / WAVE_BLOCK);
    k_generate<MULTI><<<gen_grid, WAVE_BLOCK, 0, stream>>>(*L.s0, *L.j0, L.scenes, L.jobs, L.units, ps, w.queue[0], w.counts);
    uint32_t* cur = w.queue[0];
    uint32_t* nxt = w.queue[1];
    uint32_t* spare = w.queue[3];
    uint32_t* shadow = w.queue[2];
    // first: the closest rays are camera rays (origin = camera position, not stored per path)
    auto trace = [&](const uint32_t* cq, const uint32_t* cc, uint32_t first, const uint32_t* sq, const uint32_t* sc) {
        EventScope tm(lg, EV_EXTEND, stream);
        if (counting)
            launch_trace<true, MULTI>(w.cfg.lds_depth, w.cfg.traversal, L.trace_grid, stream, *L.s0, L.scenes, L.jobs, ps,
                                      cq, cc, first, sq, sc, w.ovf, ctr);
        else
            launch_trace<false, MULTI>(w.cfg.lds_depth, w.cfg.traversal, L.trace_grid, stream, *L.s0, L.scenes, L.jobs, ps,
                                       cq, cc, first, sq, sc, w.ovf, ctr);
    };
    for (uint32_t b = 0; b <= max_bounces; ++b) {
        uint32_t* cnt = w.counts + 2 * b;
        if (w.cfg.sort_key == SORT_OCTANT && b > 0) {
            sort_queue<MULTI>(w, L, SORT_OCTANT, cur, cnt, spare);
            std::swap(cur, spare);
        }
        trace(cur, cnt, b == 0 ? 1u : 0u, b > 0 ? shadow : nullptr, b > 0 ? cnt - 1 : nullptr);
        if (w.cfg.sort_key == SORT_MATERIAL) {
            sort_queue<MULTI>(w, L, SORT_MATERIAL, cur, cnt, spare);
            std::swap(cur, spare);
        }
        {
            EventScope tm(lg, EV_SHADE, stream);
#define PT_SHADE_LAUNCH(A, W)                                                                                   \
    k_shade<MULTI, A, W><<<L.shade_grid, WAVE_BLOCK, 0, stream>>>(*L.s0, L.scenes, L.jobs, *L.j0, ps, b, cur, cnt, nxt, \
                                                                  w.counts + 2 * (b + 1), shadow, cnt + 1, w.ovf)
            // automatic window: 4 rounds for a launch over several jobs, 2 for one
            const uint32_t win = w.cfg.shade_window ? w.cfg.shade_window : (MULTI ? 4u : 2u);
            switch ((L.any_alpha ? 8u : 0u) | win) {
                case 1: PT_SHADE_LAUNCH(false, 1); break;
                case 4: PT_SHADE_LAUNCH(false, 4); break;
                case 8 | 1: PT_SHADE_LAUNCH(true, 1); break;
                case 8 | 4: PT_SHADE_LAUNCH(true, 4); break;
                case 8 | 2: PT_SHADE_LAUNCH(true, 2); break;
                default: PT_SHADE_LAUNCH(false, 2); break;
            }
#undef PT_SHADE_LAUNCH
        }
        std::swap(cur, nxt);
    }
    trace(nullptr, nullptr, 0u, shadow, w.counts + 2 * max_bounces + 1);
    const uint32_t acc_grid = std::min<uint32_t>(grid, (L.units.n_units * 64u + WAVE_BLOCK - 1) / WAVE_BLOCK);
    k_accumulate<MULTI><<<acc_grid, WAVE_BLOCK, 0, stream>>>(*L.j0, L.jobs, L.units, ps);
}

}  // namespace

// Timing is best effort: an event that cannot be created or recorded leaves
// its pair out of the sums instead of failing the render.
void EventLog::take(hipEvent_t& a, hipEvent_t& b) {
    if (pool.size() < 2) {
        a = b = nullptr;
        if (hipEventCreate(&a) != hipSuccess) a = nullptr;
        if (hipEventCreate(&b) != hipSuccess) b = nullptr;
        return;
    }
    a = pool.back();
    pool.pop_back();
    b = pool.back();
    pool.pop_back();
}
void EventLog::reset() {
    for (auto& p : pairs) {
        if (p.a) pool.push_back(p.a);
        if (p.b) pool.push_back(p.b);
    }
    pairs.clear();
}
void EventLog::sum(float ms[EV_KINDS], uint32_t launches[EV_KINDS]) {
    for (int k = 0; k < EV_KINDS; ++k) ms[k] = 0.0f, launches[k] = 0;
    for (auto& p : pairs) {
        float t = 0.0f;
        if (!p.a || !p.b || hipEventSynchronize(p.b) != hipSuccess || hipEventElapsedTime(&t, p.a, p.b) != hipSuccess)
            continue;
        ms[p.kind] += t;
        launches[p.kind] += 1;
    }
}
float EventLog::busy(int kind) {
    // intervals relative to the first usable pair's start event
    hipEvent_t ref = nullptr;
    std::vector<std::pair<float, float>> iv;
    for (auto& p : pairs) {
        if (p.kind != kind || !p.a || !p.b || hipEventSynchronize(p.b) != hipSuccess) continue;
        if (!ref) ref = p.a;
        float t0 = 0.0f, t1 = 0.0f;
        if (hipEventElapsedTime(&t0, ref, p.a) != hipSuccess || hipEventElapsedTime(&t1, ref, p.b) != hipSuccess) continue;
        iv.push_back({t0, t1});
    }
    std::sort(iv.begin(), iv.end());
    float total = 0.0f, lo = 0.0f, hi = 0.0f;
    bool open = false;
    for (auto& x : iv) {
        if (open && x.first <= hi) {
            hi = std::max(hi, x.second);
            continue;
        }
        if (open) total += hi - lo;
        lo = x.first;
        hi = x.second;
        open = true;
    }
    if (open) total += hi - lo;
    return total;
}
EventLog::~EventLog() {
    reset();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// Teardown has no caller to report to: failures of the releases are ignored.
Wavefront::~Wavefront() {
    if (partner) {
        (void)hipStreamSynchronize(partner->stream_override);
        delete partner;
    }
    if (arena) (void)hipFree(arena);
    if (ovf) (void)hipFree(ovf);
    if (tables) (void)hipFree(tables);
    if (staging) (void)hipHostFree(staging);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_done) (void)hipEventDestroy(ev_done);
}

pt_status Wavefront::ensure_overflow(size_t need, hipStream_t stream) {
    if (need <= ovf_bytes) return PT_OK;
    if (ovf) {
        hipError_t se = hipStreamSynchronize(stream);
        if (se != hipSuccess) return hip_error(se, "wavefront overflow stacks");
        (void)hipFree(ovf);
        ovf = nullptr;
    }
    hipError_t e = hipMalloc(&ovf, need);
    if (e != hipSuccess) {
        ovf_bytes = 0;
        return hip_error(e, "wavefront overflow stacks");
    }
    ovf_bytes = need;
    return PT_OK;
}

pt_status Wavefront::reserve(uint32_t paths) {
    if (paths <= capacity && arena) return PT_OK;
    if (arena) {
        hipError_t se = hipStreamSynchronize(stream_override ? stream_override : ctx->stream);
        if (se != hipSuccess) return hip_error(se, "wavefront path state");
        (void)hipFree(arena);
        arena = nullptr;
        capacity = 0;
    }
    // PATH_FIELDS per-path 4-byte arrays + 4 queues, each 256-byte aligned, + counters + sort histogram
    const size_t field_bytes = ((size_t)paths * 4 + 255) & ~(size_t)255;
    arena_bytes = (PATH_FIELDS + 4) * field_bytes + ((MAX_COUNTS * 4 + 255) & ~(size_t)255) + SORT_BUCKETS * 4;
    hipError_t e = hipMalloc(&arena, arena_bytes);
    if (e != hipSuccess) {
        arena = nullptr;
        return hip_error(e, "wavefront path state");
    }
    capacity = paths;
    char* p = (char*)arena;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 255) & ~(size_t)255;
        return r;
    };
    float** fields[] = {&ps.ox, &ps.oy, &ps.oz, &ps.dx, &ps.dy, &ps.dz, &ps.tr, &ps.tg, &ps.tb, &ps.lr, &ps.lg, &ps.lb,
                        &ps.prev_pdf, &ps.ht, &ps.hv, &ps.hw, &ps.sox, &ps.soy, &ps.soz, &ps.sdx, &ps.sdy, &ps.sdz,
                        &ps.sdist, &ps.scr, &ps.scg, &ps.scb};
    for (float** fp : fields) *fp = (float*)take(field_bytes);
    ps.rng = (uint32_t*)take(field_bytes);
    ps.job = (uint32_t*)take(field_bytes);
    ps.hit = (int32_t*)take(field_bytes);
    ps.slight = (int32_t*)take(field_bytes);
    static_assert(sizeof(fields) / sizeof(fields[0]) + 4 == PATH_FIELDS, "PathState field count");
    for (auto& q : queue) q = (uint32_t*)take(field_bytes);
    counts = (uint32_t*)take(MAX_COUNTS * 4);
    sort_hist = (uint32_t*)take(SORT_BUCKETS * 4);
    if ((size_t)(p - (char*)arena) > arena_bytes) return set_error(PT_ERR_STATE, "path arena overrun");
    return PT_OK;
}

pt_status Wavefront::render(const std::vector<HostJob>& jobs, int profiling, EventLog* log,
                            unsigned long long* host_counters) {
    if (jobs.empty()) return PT_OK;
    cfg = ctx->wave_cfg;
    const hipStream_t stream = stream_override ? stream_override : ctx->stream;
    const bool multi = jobs.size() > 1;
    // work units: (job, 64-pixel block), all of the job's samples
    uint32_t max_unit_paths = 0, max_bounces = 0, max_depth = 1;
    bool any_alpha = false;
    size_t total_units = 0;
    for (const HostJob& h : jobs) {
        const WaveJob& j = h.job;
        if (j.samples == 0 || j.tile_w == 0 || j.tile_h == 0) return set_error(PT_ERR_INVALID, "empty job");
        if ((uint64_t)j.samples * 64u > (1u << 30)) return set_error(PT_ERR_INVALID, "too many samples in one job");
        max_unit_paths = std::max(max_unit_paths, 64u * j.samples);
        max_bounces = std::max(max_bounces, j.max_bounces);
        max_depth = std::max(max_depth, h.scene->bvh_depth + 1u);
        any_alpha = any_alpha || !h.scene->view.all_opaque;
        const uint32_t be = h.block_end ? h.block_end : tile_blocks(j);
        if (h.block_begin > be || be > tile_blocks(j)) return set_error(PT_ERR_INVALID, "job block range out of the tile");
        total_units += be - h.block_begin;
    }
    uint64_t total_paths = 0;
    for (const HostJob& h : jobs) {
        const uint32_t be = h.block_end ? h.block_end : tile_blocks(h.job);
        total_paths += (uint64_t)(be - h.block_begin) * 64u * h.job.samples;
    }
    uint32_t chunk = cfg.chunk_paths;
    if (chunk == 0) {
        chunk = CHUNK_DEFAULT;
        size_t free_bytes = 0, total_bytes = 0;
        if (total_paths >= CHUNK_LARGE_MIN_PATHS && hipMemGetInfo(&free_bytes, &total_bytes) == hipSuccess &&
            free_bytes >= CHUNK_LARGE_MIN_FREE)
            chunk = CHUNK_LARGE;
    }
    uint32_t chunk_cap = std::max(chunk, max_unit_paths);
    // lanes: a chunk's last, small bounces run beside another chunk's first,
    // large ones, and shading beside traversal. Two lanes on request, or
    // automatically for a render of three chunks or more, or of at least
    // LANE_SPLIT_PATHS paths in one or two chunks, which is then cut into two
    // equal chunks (a 1080p 8-spp job: 2.14 -> 2.20 G path samples/s).
    const uint64_t LANE_SPLIT_PATHS = cfg.lane_split_paths;
    const bool split = !one_lane && total_paths <= 2ull * chunk_cap &&
                       (cfg.streams == 2 || (cfg.streams == 0 && total_paths >= LANE_SPLIT_PATHS));
    const uint32_t lanes =
        !one_lane && (cfg.streams == 2 || (cfg.streams == 0 && (split || total_paths > 2ull * chunk_cap))) ? 2u : 1u;
    if (split) chunk_cap = std::max<uint32_t>(max_unit_paths, (uint32_t)((total_paths + 1) / 2));
    pt_status st = reserve(chunk_cap);
    if (st) return st;
    const uint32_t grid = ctx->cu_count * 8u;
    const uint32_t trace_grid = ctx->cu_count * cfg.trace_waves_per_cu / (TRACE_BLOCK / 64u);
    const uint32_t shade_grid = ctx->cu_count * cfg.shade_blocks_per_cu;
    const uint32_t lds_min = std::min(cfg.lds_depth, MEGA_LDS_DEPTH);
    const size_t need = (size_t)(max_depth > lds_min ? max_depth - lds_min : 1u) *
                        std::max<size_t>((size_t)shade_grid * WAVE_BLOCK, (size_t)trace_grid * TRACE_BLOCK) * 4u;
    if ((st = ensure_overflow(need, stream))) return st;
    Wavefront* lane[2] = {this, nullptr};
    if (lanes == 2) {
        if (!partner) {
            partner = new (std::nothrow) Wavefront();
            if (!partner) return set_error(PT_ERR_MEMORY, "out of host memory");
            partner->ctx = ctx;
            partner->stream_override = ctx->lane_stream;
            hipError_t e = hipEventCreateWithFlags(&ev_start, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_done, hipEventDisableTiming);
            if (e != hipSuccess) return hip_error(e, "second wavefront lane");
        }
        partner->cfg = cfg;
        if ((st = partner->reserve(chunk_cap)) || (st = partner->ensure_overflow(need, partner->stream_override))) return st;
        lane[1] = partner;
    }
    // host tables: scenes, jobs, then the units of every chunk (job, block, offset)
    std::vector<DevScene> scene_tab;
    std::vector<WaveJob> job_tab(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        const Scene* sc = jobs[i].scene;
        uint32_t idx = 0;
        for (; idx < scene_tab.size(); ++idx)
            if (std::memcmp(&scene_tab[idx], &sc->view, sizeof(DevScene)) == 0) break;
        if (idx == scene_tab.size()) scene_tab.push_back(sc->view);
        job_tab[i] = jobs[i].job;
        job_tab[i].scene = idx;
    }
    struct ChunkDesc {
        size_t unit_begin, unit_end;
        uint32_t paths;
    };
    std::vector<ChunkDesc> chunks;
    std::vector<uint32_t> ujob, ublock, uoff;
    ujob.reserve(total_units);
    ublock.reserve(total_units);
    uoff.reserve(total_units + 64);
    {
        ChunkDesc cur{0, 0, 0};
        for (size_t ji = 0; ji < jobs.size(); ++ji) {
            const WaveJob& j = job_tab[ji];
            const uint32_t be = jobs[ji].block_end ? jobs[ji].block_end : tile_blocks(j);
            const uint32_t up = 64u * j.samples;
            for (uint32_t b = jobs[ji].block_begin; b < be; ++b) {
                if (cur.paths + up > chunk_cap) {
                    cur.unit_end = ujob.size();
                    chunks.push_back(cur);
                    cur = ChunkDesc{ujob.size(), 0, 0};
                }
                ujob.push_back((uint32_t)ji);
                ublock.push_back(b);
                uoff.push_back(cur.paths);
                cur.paths += up;
            }
        }
        cur.unit_end = ujob.size();
        if (cur.paths) chunks.push_back(cur);
    }
    // offsets array per chunk needs its terminating total: lay out as
    // [units of chunk 0 offsets..., total0, units of chunk 1 offsets..., total1, ...]
    std::vector<uint32_t> off_packed;
    off_packed.reserve(uoff.size() + chunks.size());
    std::vector<size_t> off_start(chunks.size());
    for (size_t c = 0; c < chunks.size(); ++c) {
        off_start[c] = off_packed.size();
        for (size_t u = chunks[c].unit_begin; u < chunks[c].unit_end; ++u) off_packed.push_back(uoff[u]);
        off_packed.push_back(chunks[c].paths);
    }
    const size_t b_scenes = scene_tab.size() * sizeof(DevScene), b_jobs = job_tab.size() * sizeof(WaveJob);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bytes = al(b_scenes) + al(b_jobs) + al(ujob.size() * 4) + al(ublock.size() * 4) + al(off_packed.size() * 4);
    // the staging buffer may still feed the previous render's table copy
    hipError_t se = hipStreamSynchronize(stream);
    if (se != hipSuccess) return hip_error(se, "wavefront job tables");
    if (bytes > tables_bytes) {
        if (tables) (void)hipFree(tables);
        if (staging) (void)hipHostFree(staging);
        tables = staging = nullptr;
        tables_bytes = staging_bytes = 0;
        hipError_t e = hipMalloc(&tables, bytes);
        if (e == hipSuccess) e = hipHostMalloc(&staging, bytes, 0);
        if (e != hipSuccess) return hip_error(e, "wavefront job tables");
        tables_bytes = staging_bytes = bytes;
    }
    char* hs = (char*)staging;
    char* ds = (char*)tables;
    size_t o = 0;
    auto put = [&](const void* src, size_t n) {
        std::memcpy(hs + o, src, n);
        char* d = ds + o;
        o += al(n);
        return d;
    };
    const DevScene* d_scenes = (const DevScene*)put(scene_tab.data(), b_scenes);
    const WaveJob* d_jobs = (const WaveJob*)put(job_tab.data(), b_jobs);
    const uint32_t* d_ujob = (const uint32_t*)put(ujob.data(), ujob.size() * 4);
    const uint32_t* d_ublock = (const uint32_t*)put(ublock.data(), ublock.size() * 4);
    const uint32_t* d_uoff = (const uint32_t*)put(off_packed.data(), off_packed.size() * 4);
    hipError_t e = hipMemcpyAsync(tables, staging, o, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_error(e, "job table upload");

    const bool counting = profiling >= 2;
    EventLog* lg = profiling >= 1 ? log : nullptr;
    unsigned long long* ctr = counting ? ctx->counters : nullptr;
    if (counting && (e = hipMemsetAsync(ctx->counters, 0, 16 * sizeof(unsigned long long), stream)) != hipSuccess)
        return hip_error(e, "traversal counters");
    EventScope total(lg, EV_TOTAL, stream);
    const bool two = lanes == 2 && chunks.size() > 1;
    if (two) {  // the second lane starts after the tables are on the device
        if ((e = hipEventRecord(ev_start, stream)) != hipSuccess ||
            (e = hipStreamWaitEvent(partner->stream_override, ev_start, 0)) != hipSuccess)
            return hip_error(e, "second lane start");
    }
    for (size_t c = 0; c < chunks.size(); ++c) {
        Wavefront& w = *lane[two ? c % 2 : 0];
        const hipStream_t lane_stream = &w == this ? stream : w.stream_override;
        Launch L;
        L.grid = grid;
        L.trace_grid = trace_grid;
        L.shade_grid = shade_grid;
        L.stream = lane_stream;
        L.s0 = &scene_tab[0];
        L.j0 = &job_tab[0];
        L.scenes = d_scenes;
        L.jobs = d_jobs;
        L.units.job = d_ujob + chunks[c].unit_begin;
        L.units.block = d_ublock + chunks[c].unit_begin;
        L.units.offset = d_uoff + off_start[c];
        L.units.n_units = (uint32_t)(chunks[c].unit_end - chunks[c].unit_begin);
        L.units.n_paths = chunks[c].paths;
        L.units.unit_paths = 64u * job_tab[ujob[chunks[c].unit_begin]].samples;
        for (size_t u = chunks[c].unit_begin; u < chunks[c].unit_end; ++u)
            if (64u * job_tab[ujob[u]].samples != L.units.unit_paths) {
                L.units.unit_paths = 0;
                break;
            }
        {
            // a job's blocks are all active when its tile is whole 8x8 blocks inside the image
            auto whole = [&](uint32_t ji) {
                const WaveJob& jb = job_tab[ji];
                return jb.tile_w % 8u == 0 && jb.tile_h % 8u == 0 && (uint64_t)jb.origin_x + jb.tile_w <= jb.image_w &&
                       (uint64_t)jb.origin_y + jb.tile_h <= jb.image_h;
            };
            const size_t ub = chunks[c].unit_begin;
            const uint32_t j0i = ujob[ub], b0 = ublock[ub];
            bool all = true, contiguous = L.units.unit_paths != 0;
            uint32_t last_job = UINT32_MAX;
            for (size_t u = ub; u < chunks[c].unit_end && (all || contiguous); ++u) {
                if (ujob[u] != last_job) {
                    last_job = ujob[u];
                    all = all && whole(last_job);
                }
                contiguous = contiguous && ujob[u] == j0i && ublock[u] == b0 + (uint32_t)(u - ub);
            }
            L.units.all_active = all ? 1u : 0u;
            L.units.contiguous = all && contiguous ? 1u : 0u;
            L.units.job0 = j0i;
            L.units.block0 = b0;
        }
        L.any_alpha = any_alpha;
        if (multi) run_chunk<true>(w, L, max_bounces, counting, lg, ctr);
        else run_chunk<false>(w, L, max_bounces, counting, lg, ctr);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "wavefront launch");
        if (counting) {
            uint32_t h[MAX_COUNTS];
            if ((e = hipMemcpyAsync(h, w.counts, sizeof(h), hipMemcpyDeviceToHost, lane_stream)) != hipSuccess ||
                (e = hipStreamSynchronize(lane_stream)) != hipSuccess)
                return hip_error(e, "ray counts");
            for (uint32_t b = 0; b <= max_bounces; ++b) {
                host_counters[0] += h[2 * b];
                host_counters[1] += h[2 * b + 1];
            }
        }
    }
    if (two) {  // the caller's stream resumes after both lanes
        if ((e = hipEventRecord(ev_done, partner->stream_override)) != hipSuccess ||
            (e = hipStreamWaitEvent(stream, ev_done, 0)) != hipSuccess)
            return hip_error(e, "second lane join");
    }
    return PT_OK;
}

}  // namespace ptcore
