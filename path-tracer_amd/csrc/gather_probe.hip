// gather_probe.hip - the vector L1's gather rate, measured on the device.
//
// The traversal kernel loads its child-pair records as 16-byte loads whose
// lanes land on scattered 64-byte segments once a wave's rays diverge; the
// rate at which the L1's data-return path serves such lane accesses is the
// ceiling that kernel runs against (not HBM bandwidth: its nodes are mostly
// L1 and L2 hits). This probe measures that rate: every lane of a one-wave
// block issues eight independent 16-byte loads per round from pseudo-random
// 64-byte segments of an L1-resident 16 KiB buffer, either each lane its own
// segment or four consecutive lanes the four parts of one segment.
#include "kernels.h"

namespace ptd {

constexpr uint32_t PROBE_SEGMENTS = 256;  // 16 KiB of 64-byte segments
constexpr uint32_t PROBE_ROUNDS = 64;
constexpr uint32_t PROBE_LOADS = 8;       // independent loads in flight per lane and round

__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(64) void k_gather_probe(const float4* __restrict__ data, uint32_t shared,
                                                     float* __restrict__ out) {
    const uint32_t lane = threadIdx.x;
    const uint32_t gid = blockIdx.x * 64u + lane;
    const uint32_t part = shared ? (lane & 3u) : 0u;
    uint32_t seed = probe_hash((shared ? gid >> 2 : gid) * 2654435761u + 12345u);
    float acc = 0.0f;
    for (uint32_t r = 0; r < PROBE_ROUNDS; ++r) {
        uint32_t idx[PROBE_LOADS];
#pragma unroll
        for (uint32_t k = 0; k < PROBE_LOADS; ++k) {
            seed = seed * 1664525u + 1013904223u;
            idx[k] = ((seed >> 7) & (PROBE_SEGMENTS - 1u)) * 4u + part;  // < PROBE_SEGMENTS * 4 float4
        }
        float v[PROBE_LOADS];
#pragma unroll
        for (uint32_t k = 0; k < PROBE_LOADS; ++k) {
            const float4 x = data[idx[k]];
            v[k] = (x.x + x.y) + (x.z + x.w);
        }
#pragma unroll
        for (uint32_t k = 0; k < PROBE_LOADS; ++k) acc += v[k];
    }
    out[gid] = acc;
}

}  // namespace ptd

namespace ptcore {

size_t gather_probe_data_bytes() { return (size_t)ptd::PROBE_SEGMENTS * 64u; }
uint64_t gather_probe_lane_accesses(uint32_t blocks) {
    return (uint64_t)blocks * 64u * ptd::PROBE_ROUNDS * ptd::PROBE_LOADS;
}
hipError_t launch_gather_probe(const float4* data, uint32_t shared, float* out, uint32_t blocks, hipStream_t stream) {
    ptd::k_gather_probe<<<blocks, 64, 0, stream>>>(data, shared, out);
    return hipGetLastError();
}

}  // namespace ptcore
