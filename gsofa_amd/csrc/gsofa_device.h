// gsofa_device.h -- wave64 / memory-model helpers shared by the traversal
// kernels (device code only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsofa {
namespace dev {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t bcast0(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t bcast0_64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

template <class T>
__device__ __forceinline__ T ld_agent(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Possibly stale read (may be served by the CU's L1).  Labels only decrease and
// markers only increase within a source's traversal, so a stale label is >= the
// true one and a stale marker <= the true one: such reads only filter work.
__device__ __forceinline__ uint32_t ld_cached(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t amin(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t amax(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Current value as seen by the atomics (performed in L2), bypassing a possibly
// stale L1 line.
__device__ __forceinline__ uint32_t ld_fresh(uint32_t* p) {
    return __hip_atomic_fetch_or(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long aadd64(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive wave64 prefix sum.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t y = __shfl_up(x, d, WAVE);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, WAVE);
    return x;
}

}  // namespace dev
}  // namespace gsofa
