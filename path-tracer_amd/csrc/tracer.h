// tracer.h - the pt_tracer object (PathTracer of path_tracer.rs:13-39).
#pragma once

#include "wavefront.h"

struct pt_tracer {
    ptcore::Context* ctx = nullptr;
    ptcore::Scene* scene = nullptr;
    uint64_t scene_version = 0;
    uint32_t tile_w = 0, tile_h = 0;
    PtTraceParameters params{};
    int schedule = PT_SCHEDULE_WAVEFRONT;
    float4* accum = nullptr;                 // tile_w * tile_h running sums
    float4* storage = nullptr;               // EXPOSURE * sum / sample_count
    int profiling = 0;                       // 0 off, 1 event timing, 2 timing + counters
    ptcore::EventLog log;                    // kernel events since set_profiling
    unsigned long long counters[6] = {0};    // summed device counters since set_profiling
    ptcore::Wavefront* wave = nullptr;
    void* scratch = nullptr;                 // display / primary-hit staging, grown on demand
    size_t scratch_bytes = 0;
};

namespace ptcore {
// Reset the tracer's accumulation if its scene changed (path_tracer.rs:272-282).
void tracer_sync_version(pt_tracer* t);
// The wavefront job of the tracer's next `samples` samples.
HostJob tracer_job(pt_tracer* t, uint32_t samples);
// PT_ERR_STATE when the tracer has no tile buffers (a failed resize freed them).
pt_status tracer_require_tile(const pt_tracer* t);
}  // namespace ptcore
