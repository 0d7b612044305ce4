// kernels.h - launch arguments and launcher entry points of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "pt_device.h"

namespace ptd {

// entries of the per-thread traversal stack kept in LDS by the megakernel and
// the query kernels; deeper entries go to the global overflow area
constexpr uint32_t MEGA_LDS_DEPTH = 16;
// bytes of the static device array that pt_render_tile_noalloc's kernels use
// for traversal overflow stacks
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;

// One tile of one image (path_tracer.rs:45-57 plus the tile size and the
// number of consecutive samples to trace in this launch).
struct TileArgs {
    uint32_t image_w, image_h;
    uint32_t origin_x, origin_y;
    uint32_t tile_w, tile_h;
    uint32_t max_bounces;
    uint32_t first_sample;  // 1-based sample index of the first sample
    uint32_t samples;       // consecutive samples per pixel in this launch
};

// Sort keys applied to the wavefront queues (autotuner choice).
enum SortKey : uint32_t { SORT_NONE = 0, SORT_OCTANT = 1, SORT_MATERIAL = 2 };

// automatic chunk size (WaveConfig::chunk_paths = 0): CHUNK_DEFAULT paths, or
// CHUNK_LARGE for a launch of at least CHUNK_LARGE_MIN_PATHS paths while the
// device has CHUNK_LARGE_MIN_FREE bytes free (the larger path state of two
// lanes, about 20 GB, stays a small part of that). Fewer, larger chunks leave
// fewer launch tails: a 4K 64-spp render runs 2.7% faster at 2^26 than at 2^24.
constexpr uint32_t CHUNK_DEFAULT = 1u << 24;
constexpr uint32_t CHUNK_LARGE = 1u << 26;
constexpr uint64_t CHUNK_LARGE_MIN_PATHS = 1ull << 28;
constexpr size_t CHUNK_LARGE_MIN_FREE = size_t(96) << 30;

// wavefront tuning (the autotuner sweeps these)
struct WaveConfig {
    uint32_t lds_depth = 16;          // traversal stack entries in LDS (16/24/32)
    uint32_t chunk_paths = 0;         // paths in flight per chunk (queue tile size); 0: automatic
    uint32_t sort_key = SORT_NONE;    // ray-queue ordering
    uint32_t traversal = 0;           // a PT_TRAVERSAL_* mode
    uint32_t streams = 0;             // chunk lanes running concurrently: 1, 2, or 0 = automatic
    uint32_t trace_waves_per_cu = 512;    // traversal grid: one-wave blocks per CU
    uint32_t shade_blocks_per_cu = 32;    // shade grid: 256-thread blocks per CU
    uint32_t shade_window = 0;            // queue rounds a shade block reorders into hits and misses: 1, 2 or 4;
                                          // 0: 4 for a launch over several jobs, 2 for one
    uint32_t lane_split_paths = 1u << 21; // smallest one- or two-chunk render that two lanes share
};
enum Traversal : uint32_t { TRAVERSAL_BATCH = 0, TRAVERSAL_REFILL = 1, TRAVERSAL_FETCH = 2 };

}  // namespace ptd

namespace ptcore {

hipError_t launch_megakernel(const ptd::DevScene& s, const ptd::TileArgs& a, float4* accum, float4* storage,
                             uint32_t* ovf, uint32_t grid, unsigned long long* counters, hipStream_t stream);
hipError_t launch_primary_hits(const ptd::DevScene& s, const ptd::TileArgs& a, int32_t* tri, float* t, uint32_t* ovf,
                               uint32_t grid, hipStream_t stream);
hipError_t launch_intersect(const ptd::DevScene& s, const float* rays, uint32_t count, int32_t* hit, float* t, float* v,
                            float* w, uint32_t* pairs, uint32_t* tris, uint32_t* ovf, uint32_t grid, hipStream_t stream);
hipError_t launch_occluded(const ptd::DevScene& s, const float* rays, const float* tmax, const int32_t* skip,
                           uint32_t count, uint8_t* out, uint32_t* ovf, uint32_t grid, hipStream_t stream);
hipError_t launch_display(const float4* storage, uint32_t n, uchar4* out, hipStream_t stream);
// Device address of the no-alloc scratch array on `device` (which must be current).
hipError_t noalloc_scratch_address(int device, uint32_t** out);
// The L1 gather-rate probe (gather_probe.hip): its data buffer size, the lane
// accesses one launch of `blocks` one-wave blocks makes, and the launch.
size_t gather_probe_data_bytes();
uint64_t gather_probe_lane_accesses(uint32_t blocks);
hipError_t launch_gather_probe(const float4* data, uint32_t shared, float* out, uint32_t blocks, hipStream_t stream);
// The restricted-domain sqrt_nt_ / div_nt_ against sqrtf and division
// (fast_math_check.hip): bit patterns lo .. lo + n - 1 (lo + n <= 2^32), and n
// seeded operand pairs; each adds its mismatch count to *mismatches.
hipError_t launch_check_sqrt(uint32_t lo, uint64_t n, unsigned long long* mismatches, uint32_t blocks,
                             hipStream_t stream);
hipError_t launch_check_div(uint64_t n, uint32_t seed, unsigned long long* mismatches, uint32_t blocks,
                            hipStream_t stream);

}  // namespace ptcore
