// pt_runtime.h - device-side runtime objects behind the C ABI handles.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "kernels.h"
#include "pt_internal.h"

namespace ptcore {

pt_status hip_error(hipError_t e, const char* what);
// Makes `device` current for the calling thread.
inline pt_status use_device(int device) {
    const hipError_t e = hipSetDevice(device);
    return e == hipSuccess ? PT_OK : hip_error(e, "hipSetDevice");
}

// One HIP device and stream (the reference's wgpu device + queue).
struct Context {
    int device = 0;
    uint32_t cu_count = 0;
    uint32_t clock_khz = 0, l2_bytes = 0;
    uint64_t hbm_bytes = 0;
    char arch[64] = {0};
    hipStream_t stream = nullptr;
    // the second stream of two-lane work: every render's second chunk lane and
    // the render server's second round lane. A process holds few streams this
    // way, and the runtime spreads them over its hardware queues (4 by
    // default), so the two lanes do not share a queue and serialise.
    hipStream_t lane_stream = nullptr;
    float* srgb_lut = nullptr;            // 256-entry sRGB decode table
    uint32_t* ovf = nullptr;              // grid-strided traversal overflow stacks
    size_t ovf_bytes = 0;
    unsigned long long* counters = nullptr;  // device counters of the last profiled launch
    uint32_t* noalloc_scratch = nullptr;     // device address of the no-alloc scratch array
    ptd::WaveConfig wave_cfg;                // tuning of every wavefront render on this context
    struct Wavefront* batch_wave = nullptr;  // path state of pt_trace_batch / autotune
    // make room for `depth` stack entries for `grid_threads` threads
    pt_status ensure_overflow(uint32_t depth, uint32_t grid_threads);
};

// A scene resident in device memory (WorldResources).
struct Scene {
    Context* ctx = nullptr;
    ptd::DevScene view{};
    const float4* tnode_alloc = nullptr;  // allocation behind view.tnode (which starts TNODE_LEAD entries in)
    uint32_t bvh_depth = 0;
    uint32_t triangle_count = 0, material_count = 0, bvh_node_count = 0;
    uint32_t scene_alias_capacity = 0, skybox_alias_count = 0;
    uint32_t alias_count = 0;  // scene alias entries currently on the device
    uint64_t footprint = 0;
    uint64_t version = 0;  // bumped by every update; tracers reset accumulation on change
    // render-server jobs over this scene that are not yet done, failed or
    // cancelled; destruction and updates are refused while nonzero
    std::atomic<uint32_t> server_jobs{0};
};
// PT_ERR_STATE when server jobs still use the scene
pt_status scene_require_idle(const Scene* s, const char* what);

// RAII device allocation for temporaries of one call.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    pt_status alloc(size_t n) {
        hipError_t e = hipMalloc(&ptr, n ? n : 1);
        if (e != hipSuccess) return hip_error(e, "temporary device buffer");
        bytes = n;
        return PT_OK;
    }
    ~DeviceBuffer() {
        if (ptr) (void)hipFree(ptr);  // teardown: nothing to report
    }
};

}  // namespace ptcore

struct pt_context : ptcore::Context {};
struct pt_scene : ptcore::Scene {};
