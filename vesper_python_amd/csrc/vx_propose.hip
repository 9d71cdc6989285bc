// Bookkeeping of the learned rotation proposal on the device: label
// z-scores, row gathers for minibatches and scored rows, the mask that
// drops already-scored rows from a new proposal, ordered compaction of the
// rows still to score, and the scatter of their exact scores.  No
// counterpart in the reference (which scores every Euler row,
// VESPER_1_prob_new.py:644-655).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <algorithm>

#include "vx_common.hpp"
#include "vx_internal.hpp"

namespace vx {

constexpr int kPropBlocks = 256;

// ---------------------------------------------------------------------------
// z-score: out = (in - mean) / (std + 1e-30), std with n - 1 in the
// denominator; FP64 sums in a fixed order (per-block partials, then one
// block adds them in block order), so repeated calls agree bit for bit.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double block_sum_d(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    return t;
}

__global__ void __launch_bounds__(256) k_zs_partial(int n, const float* __restrict__ in, const double* __restrict__ mean,
                                                    double* __restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    const double m = mean ? *mean : 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double d = (double)in[i] - m;
        s += mean ? d * d : d;
    }
    const double t = block_sum_d(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// stats[0] = sum(part) / n (mode 0), stats[1] = sqrt(sum(part) / (n - 1)) + 1e-30 (mode 1)
__global__ void k_zs_finish(int n, const double* __restrict__ part, int nb, int mode, double* __restrict__ stats) {
    if (threadIdx.x != 0) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += part[b];
    if (mode == 0) stats[0] = t / n;
    else stats[1] = sqrt(t / (n > 1 ? n - 1 : 1)) + 1e-30;
}

__global__ void k_zs_apply(int n, const float* __restrict__ in, const double* __restrict__ stats,
                           float* __restrict__ out) {
    const double m = stats[0], s = stats[1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (float)(((double)in[i] - m) / s);
}

// ---------------------------------------------------------------------------
// Row gather: dst row r = src row idx[r] for r < nvalid, zeros for nvalid <=
// r < nout (row_bytes a multiple of 4).
// ---------------------------------------------------------------------------

__global__ void k_gather_rows(int nout, int nvalid, const int32_t* __restrict__ idx, const uint32_t* __restrict__ src,
                              int words, uint32_t* __restrict__ dst) {
    const long long total = (long long)nout * words;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(e / words), w = (int)(e - (long long)r * words);
        dst[e] = r < nvalid ? src[(size_t)idx[r] * words + w] : 0u;
    }
}

// ---------------------------------------------------------------------------
// Scored rows fall just below the lowest prediction of the unscored ones
// (finite, so a histogram selection keeps its resolution).
// ---------------------------------------------------------------------------

__global__ void k_min_unscored(int n, const float* __restrict__ v, const uint8_t* __restrict__ scored,
                               unsigned* __restrict__ mn) {
    unsigned lo = 0xffffffffu;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (!scored[i]) lo = min(lo, ordered_bits(v[i]));
    for (int off = 32; off > 0; off >>= 1) lo = min(lo, (unsigned)__shfl_xor((int)lo, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMin(mn, lo);
}

__global__ void k_set_scored(int n, float* __restrict__ v, const uint8_t* __restrict__ scored,
                             const unsigned* __restrict__ mn) {
    if (*mn == 0xffffffffu) return;     // every row scored
    const float low = unordered_bits(*mn);
    const float put = low - 1e-3f * fabsf(low) - 1e-3f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (scored[i]) v[i] = put;
}

__global__ void k_init_u32(unsigned* p, unsigned v) {
    if (threadIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Ordered compaction: out = [value(i) for i in range(n) if keep(i)], in
// order.  Pass 1 counts per block (contiguous chunks), pass 2 scans the
// counts, pass 3 writes each chunk in order (wave ballots + LDS prefix).
// ---------------------------------------------------------------------------

struct KeepUnscored {
    const int32_t* rows;
    const uint8_t* scored;
    __device__ bool keep(int i) const { return !scored[rows[i]]; }
    __device__ int32_t value(int i) const { return rows[i]; }
};

struct KeepTopBins {
    const float* v;
    const unsigned* mm;   // ordered bits of min, max
    int bin;
    __device__ bool keep(int i) const {
        const float lo = unordered_bits(mm[0]), hi = unordered_bits(mm[1]);
        const float scale = hi > lo ? 4096.f / (hi - lo) : 0.f;
        int b = (int)((v[i] - lo) * scale);
        b = b < 0 ? 0 : (b > 4095 ? 4095 : b);
        return b >= bin;
    }
    __device__ int32_t value(int i) const { return i; }
};

template <class P>
__global__ void __launch_bounds__(256) k_cmp_count(int n, int chunk, P p, unsigned* __restrict__ cnt) {
    __shared__ unsigned s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    unsigned c = 0;
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) c += p.keep(i) ? 1u : 0u;
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = s;
}

__global__ void k_cmp_scan(int nb, unsigned* __restrict__ cnt, unsigned* __restrict__ total) {
    if (threadIdx.x != 0) return;
    unsigned acc = 0;
    for (int b = 0; b < nb; ++b) {
        const unsigned c = cnt[b];
        cnt[b] = acc;
        acc += c;
    }
    *total = acc;
}

template <class P>
__global__ void __launch_bounds__(256) k_cmp_write(int n, int chunk, P p, const unsigned* __restrict__ off,
                                                   int32_t* __restrict__ out) {
    __shared__ unsigned wsum[4];
    const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    unsigned base = off[blockIdx.x];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int t0 = b0; t0 < b1; t0 += blockDim.x) {
        const int i = t0 + threadIdx.x;
        const bool k = i < b1 && p.keep(i);
        const unsigned long long m = __ballot(k);
        const unsigned before = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) wsum[wid] = __popcll(m);
        __syncthreads();
        unsigned woff = 0;
        for (int w = 0; w < wid; ++w) woff += wsum[w];
        if (k) out[base + woff + before] = p.value(i);
        unsigned tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += wsum[w];
        base += tot;
    }
}

template <class P>
int compact_ordered(int n, P p, int32_t* out, unsigned* scratch /* >= kPropBlocks + 1 */, hipStream_t st) {
    const int nb = std::min(kPropBlocks, std::max(1, (n + 255) / 256));
    const int chunk = (n + nb - 1) / nb;
    hipLaunchKernelGGL(k_cmp_count<P>, dim3(nb), dim3(256), 0, st, n, chunk, p, scratch);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(64), 0, st, nb, scratch, scratch + kPropBlocks);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cmp_write<P>, dim3(nb), dim3(256), 0, st, n, chunk, p, scratch, out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

int compact_top_bins(int n, const float* v, const unsigned* mm, int bin, int32_t* out, unsigned* scratch,
                     hipStream_t st) {
    return compact_ordered(n, KeepTopBins{v, mm, bin}, out, scratch, st);
}

__global__ void k_scatter_scores(int n, const int32_t* __restrict__ idx, const float* __restrict__ src, int stride,
                                 float* __restrict__ dst, uint8_t* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[idx[i]] = src[(size_t)i * stride];
    if (flag) flag[idx[i]] = 1;
}

__global__ void k_fill_f32(int n, float v, float* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

}  // namespace vx

using namespace vx;

extern "C" int vx_zscore_f32(int n, const float* in, double* stats, int use_given, float* out, vx_stream stream) {
    VX_REQUIRE(n > 0 && in && stats && out, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    if (!use_given) {
        double* part = nullptr;
        VX_HIP(hipMallocAsync((void**)&part, kPropBlocks * sizeof(double), st));
        hipLaunchKernelGGL(k_zs_partial, dim3(kPropBlocks), dim3(256), 0, st, n, in, (const double*)nullptr, part);
        hipLaunchKernelGGL(k_zs_finish, dim3(1), dim3(64), 0, st, n, part, kPropBlocks, 0, stats);
        hipLaunchKernelGGL(k_zs_partial, dim3(kPropBlocks), dim3(256), 0, st, n, in, (const double*)stats, part);
        hipLaunchKernelGGL(k_zs_finish, dim3(1), dim3(64), 0, st, n, part, kPropBlocks, 1, stats);
        VX_LAUNCH_CHECK();
        VX_HIP(hipFreeAsync(part, st));
    }
    hipLaunchKernelGGL(k_zs_apply, dim3(kPropBlocks), dim3(256), 0, st, n, in, stats, out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_gather_rows(int nout, int nvalid, const int32_t* idx, const void* src, int row_bytes, void* dst,
                              vx_stream stream) {
    VX_REQUIRE(nout >= 0 && nvalid >= 0 && nvalid <= nout && dst && row_bytes > 0 && row_bytes % 4 == 0,
               VX_ERR_ARG, "bad arguments");
    VX_REQUIRE(nvalid == 0 || (idx && src), VX_ERR_ARG, "NULL idx or src");
    if (nout == 0) return VX_OK;
    const long long total = (long long)nout * (row_bytes / 4);
    const unsigned g = (unsigned)std::min<long long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_gather_rows, dim3(g), dim3(256), 0, (hipStream_t)stream, nout, nvalid, idx,
                       (const uint32_t*)src, row_bytes / 4, (uint32_t*)dst);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_mask_scored(int n, float* pred, const uint8_t* scored, vx_stream stream) {
    VX_REQUIRE(n > 0 && pred && scored, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    unsigned* mn = nullptr;
    VX_HIP(hipMallocAsync((void**)&mn, sizeof(unsigned), st));
    hipLaunchKernelGGL(k_init_u32, dim3(1), dim3(64), 0, st, mn, 0xffffffffu);
    hipLaunchKernelGGL(k_min_unscored, dim3(kPropBlocks), dim3(256), 0, st, n, pred, scored, mn);
    hipLaunchKernelGGL(k_set_scored, dim3(kPropBlocks), dim3(256), 0, st, n, pred, scored, mn);
    VX_LAUNCH_CHECK();
    VX_HIP(hipFreeAsync(mn, st));
    return VX_OK;
}

extern "C" int vx_filter_unscored(int n, const int32_t* rows, const uint8_t* scored, int32_t* out, int* count_host,
                                  vx_stream stream) {
    VX_REQUIRE(n >= 0 && count_host && (n == 0 || (rows && scored && out)), VX_ERR_ARG, "bad arguments");
    *count_host = 0;
    if (n == 0) return VX_OK;
    hipStream_t st = (hipStream_t)stream;
    unsigned* scr = nullptr;
    VX_HIP(hipMallocAsync((void**)&scr, (kPropBlocks + 1) * sizeof(unsigned), st));
    int rc = compact_ordered(n, KeepUnscored{rows, scored}, out, scr, st);
    if (rc != VX_OK) return rc;
    unsigned cnt = 0;
    VX_HIP(hipMemcpyAsync(&cnt, scr + kPropBlocks, sizeof(cnt), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    VX_HIP(hipFreeAsync(scr, st));
    *count_host = (int)cnt;
    return VX_OK;
}

extern "C" int vx_scatter_scores(int n, const int32_t* idx, const float* src, int stride, float* dst, uint8_t* flag,
                                 vx_stream stream) {
    VX_REQUIRE(n >= 0 && stride >= 1 && (n == 0 || (idx && src && dst)), VX_ERR_ARG, "bad arguments");
    if (n == 0) return VX_OK;
    hipLaunchKernelGGL(k_scatter_scores, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, idx, src,
                       stride, dst, flag);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_fill_f32(int n, float v, float* dst, vx_stream stream) {
    VX_REQUIRE(n >= 0 && (n == 0 || dst), VX_ERR_ARG, "bad arguments");
    if (n == 0) return VX_OK;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min(4096, (n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, v, dst);
    VX_LAUNCH_CHECK();
    return VX_OK;
}
