// fast_math_check.hip - the device's restricted-domain square root and
// division (sqrt_nt_, div_nt_) compared with its correctly rounded sqrtf and
// division, on the device itself.
//
// The shading code calls sqrt_nt_ and div_nt_ only where the argument is known
// to lie in their domain. The square root is checked on every float32 bit
// pattern of a range (the whole domain is 1.9e9 patterns: milliseconds), the
// division on pseudo-random operand pairs drawn from its domain: the divisor's
// magnitude in [2^-60, 2^60), the dividend's +0 or in [2^-60, 2^4), both
// signs, every mantissa equally likely.
#include "kernels.h"

namespace ptd {

__device__ __forceinline__ uint32_t check_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// same value, or both NaN
__device__ __forceinline__ bool same_result(float a, float b) {
    return ptm::float_to_bits(a) == ptm::float_to_bits(b) || (a != a && b != b);
}

// one atomic per wave that found mismatches (the lanes still in the loop vote)
__device__ __forceinline__ void count_wave(bool bad, unsigned long long* counter) {
    const uint64_t m = __ballot(bad);
    if (m != 0u && (threadIdx.x & 63u) == (uint32_t)__ffsll((unsigned long long)m) - 1u)
        atomicAdd(counter, (unsigned long long)__popcll(m));
}

// every bit pattern lo + i, i < n (lo + n <= 2^32)
__global__ __launch_bounds__(256) void k_check_sqrt(uint32_t lo, uint64_t n, unsigned long long* mismatches) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float x = ptm::bits_to_float(lo + (uint32_t)i);
        count_wave(!same_result(ptm::sqrt_nt_(x), sqrtf(x)), mismatches);
    }
}

__device__ __forceinline__ float magnitude(uint32_t h, int lo_exp, int span) {
    const int e = lo_exp + (int)((h >> 23) % (uint32_t)span);  // exponent in [lo_exp, lo_exp + span)
    return ptm::bits_to_float(((uint32_t)(e + 127) << 23) | (h & 0x7fffffu));
}

__global__ __launch_bounds__(256) void k_check_div(uint64_t n, uint32_t seed, unsigned long long* mismatches) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t h1 = check_hash((uint32_t)i * 2654435761u ^ seed ^ (uint32_t)(i >> 32) * 0x9e3779b9u);
        const uint32_t h2 = check_hash(h1 + 0x68bc21ebu);
        const uint32_t h3 = check_hash(h2 ^ 0x02e5be93u);
        float b = magnitude(h1, -60, 120);
        if (h3 & 1u) b = -b;
        float a = (h3 & 0x3f0u) == 0u ? 0.0f : magnitude(h2, -60, 64);  // +0 one time in 64
        if ((h3 & 2u) && a != 0.0f) a = -a;
        count_wave(!same_result(ptm::div_nt_(a, b), a / b), mismatches);
    }
}

}  // namespace ptd

namespace ptcore {

hipError_t launch_check_sqrt(uint32_t lo, uint64_t n, unsigned long long* mismatches, uint32_t blocks,
                             hipStream_t stream) {
    ptd::k_check_sqrt<<<blocks, 256, 0, stream>>>(lo, n, mismatches);
    return hipGetLastError();
}
hipError_t launch_check_div(uint64_t n, uint32_t seed, unsigned long long* mismatches, uint32_t blocks,
                            hipStream_t stream) {
    ptd::k_check_div<<<blocks, 256, 0, stream>>>(n, seed, mismatches);
    return hipGetLastError();
}

}  // namespace ptcore
