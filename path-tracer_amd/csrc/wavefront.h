// wavefront.h - the wavefront schedule: generate, extend, shade and shadow
// kernels over compacted path queues, then an in-order per-pixel accumulate,
// driven by a table of render jobs (one job for a tracer tile; many jobs of
// different scenes, sizes and sample counts for a batch).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"
#include "pt_runtime.h"

namespace ptd {

// Structure-of-arrays path state for `capacity` paths of one chunk.
struct PathState {
    float* ox; float* oy; float* oz;      // ray origin
    float* dx; float* dy; float* dz;      // ray direction
    float* tr; float* tg; float* tb;      // throughput
    float* lr; float* lg; float* lb;      // radiance so far (final_color)
    float* prev_pdf;                      // previous_bsdf_pdf
    uint32_t* rng;                        // RNG state
    uint32_t* job;                        // job index of the path
    int32_t* hit;                         // closest hit triangle (-1: miss)
    float* ht; float* hv; float* hw;      // hit t and barycentrics
    // pending shadow ray and its contribution if unoccluded
    float* sox; float* soy; float* soz;
    float* sdx; float* sdy; float* sdz;
    float* sdist;
    int32_t* slight;
    float* scr; float* scg; float* scb;
};
constexpr uint32_t PATH_FIELDS = 30;  // 4-byte arrays in PathState

// One render job: `samples` consecutive samples of a tile of an image of one
// scene (path_tracer.rs:45-57 plus the tile size).
struct WaveJob {
    uint32_t scene;                // index into the scene table
    uint32_t image_w, image_h, origin_x, origin_y, tile_w, tile_h;
    uint32_t max_bounces, first_sample, samples;
    float4* accum;                 // tile_w * tile_h running sums
    float4* storage;               // EXPOSURE * sum / samples (may be null)
};

// Work units of a chunk: unit u covers the 64 pixels of 8x8 block
// unit_block[u] of job unit_job[u], all of that job's samples, as paths
// unit_offset[u] + k * 64 + lane (k = sample, lane = pixel in the block).
struct ChunkUnits {
    const uint32_t* job;
    const uint32_t* block;
    const uint32_t* offset;  // n_units + 1 entries, relative to the chunk
    uint32_t n_units;
    uint32_t n_paths;
    uint32_t unit_paths;     // paths of every unit when they are all alike (one job, or equal spp), else 0
    // nonzero when every unit's pixels lie in its job's tile and image: every
    // path is then active and the bounce-0 queue is 0 .. n_paths-1
    uint32_t all_active;
    // nonzero when, besides, unit u is block block0 + u of job job0 for every
    // u, so the unit tables need not be read
    uint32_t contiguous;
    uint32_t job0, block0;
};

}  // namespace ptd

namespace ptcore {

// kernel classes timed by the event log
enum { EV_EXTEND = 0, EV_SHADE = 1, EV_SHADOW = 2, EV_TOTAL = 3, EV_KINDS = 4 };

// Pool of HIP event pairs recorded around kernels; read after the stream drains.
struct EventLog {
    struct Pair {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pair> pairs;
    std::vector<hipEvent_t> pool;
    unsigned kinds = ~0u;  // bit k: record kernels of class k (each pair is a marker on the stream)
    bool records(int kind) const { return (kinds >> kind) & 1u; }
    void take(hipEvent_t& a, hipEvent_t& b);
    void reset();  // recycle recorded pairs
    void sum(float ms[EV_KINDS], uint32_t launches[EV_KINDS]);
    // time covered by at least one interval of class `kind` (the union of its pairs)
    float busy(int kind);
    ~EventLog();
};

// A job as the host schedules it: the 8x8 pixel blocks [block_begin,
// block_end) of the job's tile (block_end 0: all blocks), each with all samples.
struct HostJob {
    Scene* scene;
    ptd::WaveJob job;
    uint32_t block_begin = 0, block_end = 0;
};
inline uint32_t tile_blocks(const ptd::WaveJob& j) { return ((j.tile_w + 7u) / 8u) * ((j.tile_h + 7u) / 8u); }

// Device counters (unsigned long long[16]) layout:
//   [0] closest rays [1] shadow rays [2] extend pair tests [3] extend tri tests
//   [4] shadow pair tests [5] shadow tri tests
struct Wavefront {
    Context* ctx = nullptr;
    uint32_t capacity = 0;      // paths per chunk
    void* arena = nullptr;      // path state + queues + counters
    size_t arena_bytes = 0;
    ptd::PathState ps{};
    uint32_t* queue[4] = {nullptr, nullptr, nullptr, nullptr};  // ray queue cur/next, shadow, sort scratch
    uint32_t* counts = nullptr;  // per bounce: [2*b] ray queue count, [2*b+1] shadow queue count
    uint32_t* sort_hist = nullptr;  // 256 bucket counters
    uint32_t* ovf = nullptr;     // overflow stacks for the traversal grid
    size_t ovf_bytes = 0;
    // job/scene/unit tables (device) and their pinned host staging
    void* tables = nullptr;
    size_t tables_bytes = 0;
    void* staging = nullptr;
    size_t staging_bytes = 0;
    ptd::WaveConfig cfg;
    hipStream_t stream_override = nullptr;  // run on this stream instead of the context's
    bool one_lane = false;                  // never split a render over a second lane (the caller pipelines)
    // second chunk lane (cfg.streams == 2): its own path state and overflow
    // stacks, on the context's lane stream; odd chunks run there, overlapping
    // the even ones
    Wavefront* partner = nullptr;
    hipEvent_t ev_start = nullptr, ev_done = nullptr;
    pt_status ensure_overflow(size_t bytes, hipStream_t stream);
    pt_status reserve(uint32_t paths);
    // Trace every job completely. profiling: 0 off, 1 per-kernel event
    // timing into `log`, 2 timing + traversal counters (synchronises per chunk).
    pt_status render(const std::vector<HostJob>& jobs, int profiling, EventLog* log, unsigned long long* host_counters);
    // device bytes held: path state, overflow stacks and tables, with the second lane's
    size_t device_bytes() const {
        return arena_bytes + ovf_bytes + tables_bytes + (partner ? partner->device_bytes() : 0);
    }
    ~Wavefront();
};

}  // namespace ptcore
