// Shared host/device helpers for libvesper_hip: error reporting across the
// C-ABI (status codes + thread-local message, never abort), ordered-float
// argmax keys, small reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "vesper_hip.h"

namespace vx {

void set_error(const char* fmt, ...);

#define VX_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::vx::set_error("%s failed: %s (%s:%d)", #call,                   \
                            hipGetErrorString(e_), __FILE__, __LINE__);       \
            return VX_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

#define VX_REQUIRE(cond, code, ...)                                           \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::vx::set_error(__VA_ARGS__);                                     \
            return (code);                                                    \
        }                                                                     \
    } while (0)

#define VX_LAUNCH_CHECK()                                                     \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess) {                                               \
            ::vx::set_error("kernel launch failed: %s (%s:%d)",               \
                            hipGetErrorString(e_), __FILE__, __LINE__);       \
            return VX_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

// Monotone map float -> uint32 (larger float -> larger key).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float unordered_bits(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}
// Argmax key: maximum value first, then the smallest flat index.
__device__ __forceinline__ unsigned long long argmax_key(float v, uint32_t idx) {
    return ((unsigned long long)ordered_bits(v) << 32) | (unsigned long long)(0xffffffffu - idx);
}

// (value, index) "better" relation of numpy's argmax: larger value, ties ->
// smaller index.
__device__ __forceinline__ bool better(float v, uint32_t i, float bv, uint32_t bi) {
    return (v > bv) || (v == bv && i < bi);
}

// Block-wide argmax; returns the winner to thread 0 via (bv, bi).
__device__ __forceinline__ void block_argmax(float& bv, uint32_t& bi, float* s_v, uint32_t* s_i) {
    for (int off = 32; off > 0; off >>= 1) {
        float ov = __shfl_xor(bv, off, 64);
        uint32_t oi = __shfl_xor(bi, off, 64);
        if (better(ov, oi, bv, bi)) { bv = ov; bi = oi; }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) { s_v[wid] = bv; s_i[wid] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; ++w)
            if (better(s_v[w], s_i[w], bv, bi)) { bv = s_v[w]; bi = s_i[w]; }
    }
}

}  // namespace vx
