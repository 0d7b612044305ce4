// wavefront.h - the wavefront schedule: generate, extend, shade and shadow
// kernels over compacted path queues, then an in-order per-pixel accumulate.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "pt_runtime.h"

namespace ptd {

// Structure-of-arrays path state for `capacity` paths. A path is one sample
// of one pixel of the chunk: path = sample * chunk_pixels + pixel.
struct PathState {
    float* ox; float* oy; float* oz;      // ray origin
    float* dx; float* dy; float* dz;      // ray direction
    float* tr; float* tg; float* tb;      // throughput
    float* lr; float* lg; float* lb;      // radiance so far (final_color)
    float* prev_pdf;                      // previous_bsdf_pdf
    uint32_t* rng;                        // RNG state
    int32_t* hit;                         // closest hit triangle (-1: miss)
    float* ht; float* hv; float* hw;      // hit t and barycentrics
    // pending shadow ray and its contribution if unoccluded
    float* sox; float* soy; float* soz;
    float* sdx; float* sdy; float* sdz;
    float* sdist;
    int32_t* slight;
    float* scr; float* scg; float* scb;
};

// Launch arguments shared by the kernels of one chunk.
struct WaveArgs {
    TileArgs tile;
    uint32_t chunk_first_pixel;  // first pixel (in tile block order) of this chunk
    uint32_t chunk_pixels;       // pixels in this chunk
    uint32_t tile_pixels;        // pixels in the tile, block order covers ceil-8 tiles
};

// extend-kernel tuning (the autotuner sweeps these)
struct ExtendConfig {
    uint32_t lds_depth = 16;  // traversal stack entries in LDS (16/24/32)
    uint32_t block = 256;     // threads per block
};

}  // namespace ptd

namespace ptcore {

struct Wavefront {
    Context* ctx = nullptr;
    uint32_t capacity = 0;      // paths per chunk
    void* arena = nullptr;      // path state + queues
    size_t arena_bytes = 0;
    ptd::PathState ps{};
    uint32_t* queue[3] = {nullptr, nullptr, nullptr};  // ray queue (cur/next) and shadow queue
    uint32_t* counts = nullptr;  // per bounce: [2*b] ray queue count, [2*b+1] shadow queue count
    uint32_t* ovf = nullptr;     // overflow stacks for the extend/shadow grid
    size_t ovf_bytes = 0;
    ptd::ExtendConfig extend_cfg;
    uint32_t max_chunk_paths = 1u << 22;  // 4M paths per chunk by default
    pt_status reserve(uint32_t paths);
    pt_status render(Scene& scene, const ptd::TileArgs& tile, float4* accum, float4* storage, int profiling,
                     pt_trace_stats* stats);
    ~Wavefront();
};

}  // namespace ptcore
