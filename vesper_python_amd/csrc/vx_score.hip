// Map statistics, final FP64 scoring (get_score), top-N, Laplacian filter,
// and the C-ABI's error/version entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <algorithm>
#include <mutex>
#include <string.h>

#include "vx_common.hpp"
#include "vx_internal.hpp"

namespace vx {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// Deterministic FP64 block reductions: K sums per block -> partials[block][K];
// a single block then adds the partials in block order.
// ---------------------------------------------------------------------------

template <int K>
__device__ __forceinline__ void block_sum_store(double (&v)[K], double* __restrict__ partial) {
    __shared__ double s[K][16];
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) s[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double t = 0.0;
            for (int w = 0; w < nw; ++w) t += s[k][w];
            partial[(size_t)blockIdx.x * K + k] = t;
        }
    }
}

constexpr int kRedBlocks = 512;

// One block of kRedBlocks threads adds the partials of kRedBlocks blocks:
// thread b holds block b's K sums, then a fixed tree (xor shuffles within
// each wave, then the waves' totals in wave order), so the order of the
// additions is the same on every call.
template <int K>
__global__ void __launch_bounds__(kRedBlocks) k_reduce_partials(const double* __restrict__ partial, int nblocks,
                                                                double* __restrict__ out) {
    __shared__ double s[K][kRedBlocks / 64];
    const int b = threadIdx.x;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = b < nblocks ? partial[(size_t)b * K + k] : 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int lane = b & 63, wid = b >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) s[k][wid] = v[k];
    __syncthreads();
    if (b < K) {
        double t = 0.0;
        for (int w = 0; w < kRedBlocks / 64; ++w) t += s[b][w];
        out[b] = t;
    }
}

// pass 1: {sum all, count != 0, count > 0, sum > 0, sum^2 > 0}
__global__ void __launch_bounds__(256) k_stats1(const double* __restrict__ d, size_t count, double* __restrict__ partial) {
    double v[5] = {0, 0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
        const double x = d[i];
        v[0] += x;
        v[1] += (x != 0.0) ? 1.0 : 0.0;
        if (x > 0.0) { v[2] += 1.0; v[3] += x; v[4] += x * x; }
    }
    block_sum_store<5>(v, partial);
}

// pass 2: sum over > 0 of (x - mean)^2, mean = totals[3] / totals[2]
__global__ void __launch_bounds__(256) k_stats2(const double* __restrict__ d, size_t count, const double* __restrict__ totals,
                                               double* __restrict__ partial) {
    const double ave = totals[3] / totals[2];
    double v[1] = {0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
        const double x = d[i];
        if (x > 0.0) { const double t = x - ave; v[0] += t * t; }
    }
    block_sum_store<1>(v, partial);
}

// get_score overlap sums at signed translation t (VESPER_1_prob_new.py:1333-1379):
// {Nm, target==0 & search>0, cc, pcc, dot, count_nonzero(target)}
__global__ void __launch_bounds__(256)
k_score(int n, const double* __restrict__ td, const float* __restrict__ tv, double ave1,
        const double* __restrict__ sd, const float* __restrict__ sv, const double* __restrict__ stot,
        int tx, int ty, int tz, double* __restrict__ partial) {
    const double ave2 = stot[3] / stot[2];
    const size_t n3 = (size_t)n * n * n;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % n), y = (int)((i / n) % n), x = (int)(i / ((size_t)n * n));
        const double d1r = td[i];
        v[5] += (d1r != 0.0) ? 1.0 : 0.0;
        const int sx = x + tx, sy = y + ty, sz = z + tz;
        if (sx < 0 || sy < 0 || sz < 0 || sx >= n || sy >= n || sz >= n) continue;
        const size_t j = ((size_t)sx * n + sy) * n + sz;
        const double ds = sd[j];
        const bool snz = ds > 0.0;
        if (d1r > 0.0 && snz) v[0] += 1.0;
        if (d1r == 0.0 && snz) v[1] += 1.0;
        const double d1 = d1r <= 0.0 ? 0.0 : d1r;       // :1356
        const double d2 = snz ? d1 : 0.0;               // :1357 (uses d1)
        const double pd1 = d1 <= 0.0 ? 0.0 : d1 - ave1; // :1359
        const double pd2 = d2 <= 0.0 ? 0.0 : d2 - ave2; // :1360
        v[2] += d1 * d2;
        v[3] += pd1 * pd2;
        v[4] += (double)tv[3 * i] * (double)sv[3 * j] + (double)tv[3 * i + 1] * (double)sv[3 * j + 1] +
                (double)tv[3 * i + 2] * (double)sv[3 * j + 2];
    }
    block_sum_store<6>(v, partial);
}

// ---------------------------------------------------------------------------
// Top-N by (score desc, row asc): bitonic sort of 64-bit keys in LDS.
// ---------------------------------------------------------------------------

constexpr int kSortN = 2048;

__device__ __forceinline__ unsigned long long row_key(float s, uint32_t row) {
    return ((unsigned long long)ordered_bits(s) << 32) | (unsigned long long)(0xffffffffu - row);
}

// Sort kSortN keys descending in LDS with 1024 threads.
__device__ void bitonic_desc(unsigned long long* k) {
    for (int size = 2; size <= kSortN; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            const int t = threadIdx.x;
            const int i = 2 * t - (t & (stride - 1));
            const int j = i + stride;
            const bool desc = ((i & size) == 0);
            const unsigned long long a = k[i], b = k[j];
            if ((a < b) == desc) { k[i] = b; k[j] = a; }
        }
    }
    __syncthreads();
}

// Pass over `n` inputs (scores with row ids, or keys from an earlier pass);
// each block sorts a chunk and writes its best `keep` keys.
__global__ void __launch_bounds__(1024)
k_topn_chunk(const float* __restrict__ scores, int stride, int col, const unsigned long long* __restrict__ in_keys,
             int n, int keep, unsigned long long* __restrict__ out_keys) {
    __shared__ unsigned long long k[kSortN];
    const int base = blockIdx.x * kSortN;
    for (int t = threadIdx.x; t < kSortN; t += blockDim.x) {
        const int i = base + t;
        unsigned long long v = 0ull;
        if (i < n) v = in_keys ? in_keys[i] : row_key(scores[(size_t)i * stride + col], (uint32_t)i);
        k[t] = v;
    }
    bitonic_desc(k);
    const int valid = min(kSortN, n - base);
    const int m = min(keep, valid);
    for (int t = threadIdx.x; t < keep; t += blockDim.x)
        out_keys[(size_t)blockIdx.x * keep + t] = (t < m) ? k[t] : 0ull;
}

__global__ void k_keys_to_rows(const unsigned long long* __restrict__ keys, int m, int32_t* __restrict__ rows) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) rows[t] = (int32_t)(0xffffffffu - (uint32_t)(keys[t] & 0xffffffffu));
}

// laplacian_filter at one voxel (VESPER_1_prob_new.py:512-525, same addition
// order): 6-neighbour stencil where the value is > 0, neighbours outside the
// grid count as 0, else 0.
__device__ __forceinline__ double laplacian_at(int n, const double* __restrict__ a, size_t v) {
    const int z = (int)(v % n), y = (int)((v / n) % n), x = (int)(v / ((size_t)n * n));
    double r = 0.0;
    const double c = a[v];
    if (c > 0) {
        r = -6.0 * c;
        if (x + 1 < n) r += a[v + (size_t)n * n];
        if (x - 1 >= 0) r += a[v - (size_t)n * n];
        if (y + 1 < n) r += a[v + n];
        if (y - 1 >= 0) r += a[v - n];
        if (z + 1 < n) r += a[v + 1];
        if (z - 1 >= 0) r += a[v - 1];
    }
    return r;
}

__global__ void k_laplacian(int n, const double* __restrict__ a, double* __restrict__ out) {
    const size_t n3 = (size_t)n * n * n;
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n3) return;
    out[v] = laplacian_at(n, a, v);
}

// The float32 correlation planes of a map for a mode: the FFT inputs of
// fft_search_best_dot's vector components (VESPER_1_prob_new.py:1434-1438,
// :1485-1497), of fft_search_score_trans_1d's mode transform (:546-558) and
// of the target channels (:601-615, :831-838).  Each value is formed in
// float64 as the reference's NumPy code does, then rounded to float32 by the
// copy into the float32 FFT input.
__global__ void __launch_bounds__(256)
k_mode_planes(int n, int mode, const double* __restrict__ data, const float* __restrict__ vec,
              const float* __restrict__ prob, double ave, float* __restrict__ out) {
    const size_t n3 = (size_t)n * n * n;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    if (mode == VX_MODE_VEC || mode == VX_MODE_PROB) {
        out[i] = vec[3 * i];
        out[n3 + i] = vec[3 * i + 1];
        out[2 * n3 + i] = vec[3 * i + 2];
        if (mode == VX_MODE_PROB)
            for (int k = 0; k < 4; ++k) out[(3 + k) * n3 + i] = prob[k * n3 + i];
        return;
    }
    const double d = data[i];
    double v = 0.0;
    if (mode == VX_MODE_LAPLACIAN) v = laplacian_at(n, data, i);
    else if (d > 0) {
        if (mode == VX_MODE_OVERLAP) v = 1.0;
        else if (mode == VX_MODE_CC) v = d;
        else if (mode == VX_MODE_PCC) v = d - ave;
    }
    out[i] = (float)v;
}

// Library-owned scratch for the reductions and top-N (grow-only, per process;
// allocated outside the timed search loop on first use).
struct Scratch {
    std::mutex mu;
    void* ptr = nullptr;
    size_t bytes = 0;
};
static Scratch g_scratch;

static int scratch_get(size_t bytes, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch.mu);
    if (g_scratch.bytes < bytes) {
        if (g_scratch.ptr) {
            VX_HIP(hipDeviceSynchronize());
            VX_HIP(hipFree(g_scratch.ptr));
            g_scratch.ptr = nullptr;
            g_scratch.bytes = 0;
        }
        VX_HIP(hipMalloc(&g_scratch.ptr, bytes));
        g_scratch.bytes = bytes;
    }
    *out = g_scratch.ptr;
    return VX_OK;
}

static int stats_device(const double* d, size_t count, double* totals /*dev: 5 + 1*/, double* partial,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_stats1, dim3(kRedBlocks), dim3(256), 0, st, d, count, partial);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials<5>, dim3(1), dim3(kRedBlocks), 0, st, partial, kRedBlocks, totals);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_stats2, dim3(kRedBlocks), dim3(256), 0, st, d, count, totals, partial);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(kRedBlocks), 0, st, partial, kRedBlocks, totals + 5);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

}  // namespace vx

using namespace vx;

extern "C" int vx_abi_version(void) { return VX_ABI_VERSION; }

extern "C" const char* vx_last_error(void) { return g_err; }

extern "C" int vx_map_stats(const double* data, int64_t count, double* out5, vx_stream stream) {
    VX_REQUIRE(data && out5 && count > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    void* scr = nullptr;
    int rc = scratch_get((size_t)kRedBlocks * 8 * sizeof(double) + 64 * sizeof(double), &scr);
    if (rc) return rc;
    double* partial = (double*)scr;
    double* totals = partial + (size_t)kRedBlocks * 8;
    rc = stats_device(data, (size_t)count, totals, partial, st);
    if (rc) return rc;
    double h[6];
    VX_HIP(hipMemcpyAsync(h, totals, sizeof(h), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    out5[0] = h[0];
    out5[1] = h[1];
    out5[2] = h[2] > 0 ? h[3] / h[2] : NAN;
    out5[3] = sqrt(h[4]);
    out5[4] = sqrt(h[5]);
    return VX_OK;
}

extern "C" int vx_get_score(int n, const double* td, const float* tv, double t_ave, const double* sd,
                            const float* sv, const int32_t* trans3, double* out, vx_stream stream) {
    VX_REQUIRE(td && tv && sd && sv && trans3 && out && n > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    void* scr = nullptr;
    int rc = scratch_get((size_t)kRedBlocks * 8 * sizeof(double) + 64 * sizeof(double), &scr);
    if (rc) return rc;
    double* partial = (double*)scr;
    double* totals = partial + (size_t)kRedBlocks * 8;   // [0..5] search stats, [8..13] overlap sums
    const size_t n3 = (size_t)n * n * n;
    rc = stats_device(sd, n3, totals, partial, st);
    if (rc) return rc;
    // translation wrap: t -= dim if t > dim / 2 (VESPER_1_prob_new.py:1323-1329)
    int t[3];
    for (int a = 0; a < 3; ++a) {
        t[a] = trans3[a];
        if ((double)trans3[a] > 0.5 * n) t[a] -= n;
    }
    hipLaunchKernelGGL(k_score, dim3(kRedBlocks), dim3(256), 0, st, n, td, tv, t_ave, sd, sv, totals, t[0], t[1],
                       t[2], partial);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials<6>, dim3(1), dim3(kRedBlocks), 0, st, partial, kRedBlocks, totals + 8);
    VX_LAUNCH_CHECK();
    double h[14];
    VX_HIP(hipMemcpyAsync(h, totals, sizeof(h), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    const double ave2 = h[2] > 0 ? h[3] / h[2] : NAN;
    out[0] = h[8];                 // Nm
    out[1] = h[13] + h[9];         // total = nnz(target) + #(target == 0 & search > 0)
    out[2] = h[10];                // cc (raw)
    out[3] = h[11];                // pcc (raw)
    out[4] = h[12];                // DOT score
    out[5] = ave2;
    out[6] = sqrt(h[4]);           // std2
    out[7] = sqrt(h[5]);           // pstd2
    out[8] = h[13];
    out[9] = h[2];
    return VX_OK;
}

namespace vx {

// Keys held by one of the two ping-pong buffers of a top-N selection: the
// first pass writes `topn` keys per chunk of kSortN rows; later passes shrink
// that list by kSortN / topn >= 2 each.
size_t topn_cap(int nrows, int topn) {
    const size_t nchunk = ((size_t)nrows + kSortN - 1) / kSortN;
    return nchunk * (size_t)topn + kSortN;
}

int topn_check(const float* scores, int nrows, int stride, int col, int topn, const int32_t* rows) {
    VX_REQUIRE(scores && rows && nrows > 0 && stride > 0 && col >= 0 && col < stride, VX_ERR_ARG, "bad arguments");
    VX_REQUIRE(topn >= 1 && topn <= 1024, VX_ERR_ARG, "topn must be in [1, 1024]");
    return VX_OK;
}

// The selection itself on caller-provided key buffers (each topn_cap keys).
int topn_run(const float* scores, int nrows, int stride, int col, int topn, int32_t* rows,
                    unsigned long long* bufA, unsigned long long* bufB, hipStream_t st) {
    const int keep = topn;
    int nb = (nrows + kSortN - 1) / kSortN;
    hipLaunchKernelGGL(k_topn_chunk, dim3(nb), dim3(1024), 0, st, scores, stride, col,
                       (const unsigned long long*)nullptr, nrows, keep, bufA);
    VX_LAUNCH_CHECK();
    int n = nb * keep;
    while (nb > 1) {
        nb = (n + kSortN - 1) / kSortN;
        hipLaunchKernelGGL(k_topn_chunk, dim3(nb), dim3(1024), 0, st, scores, stride, col, bufA, n, keep, bufB);
        VX_LAUNCH_CHECK();
        unsigned long long* tmp = bufA; bufA = bufB; bufB = tmp;
        n = nb * keep;
    }
    const int m = topn < nrows ? topn : nrows;
    hipLaunchKernelGGL(k_keys_to_rows, dim3((m + 255) / 256), dim3(256), 0, st, bufA, m, rows);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

}  // namespace vx

extern "C" int vx_topn(const float* scores, int nrows, int stride, int col, int topn, int32_t* rows,
                       vx_stream stream) {
    int rc = topn_check(scores, nrows, stride, col, topn, rows);
    if (rc) return rc;
    const size_t cap = topn_cap(nrows, topn);
    void* scr = nullptr;
    rc = scratch_get(2 * cap * sizeof(unsigned long long) + (size_t)kRedBlocks * 8 * sizeof(double) + 512, &scr);
    if (rc) return rc;
    unsigned long long* bufA = (unsigned long long*)((char*)scr + (size_t)kRedBlocks * 8 * sizeof(double) + 512);
    return topn_run(scores, nrows, stride, col, topn, rows, bufA, bufA + cap, (hipStream_t)stream);
}

extern "C" int vx_topn_ws_bytes(int nrows, int topn, int64_t* bytes_host) {
    VX_REQUIRE(bytes_host && nrows > 0 && topn >= 1 && topn <= 1024, VX_ERR_ARG, "bad arguments");
    *bytes_host = (int64_t)(2 * topn_cap(nrows, topn) * sizeof(unsigned long long));
    return VX_OK;
}

extern "C" int vx_topn_ws(const float* scores, int nrows, int stride, int col, int topn, int32_t* rows,
                          void* ws, int64_t ws_bytes, vx_stream stream) {
    int rc = topn_check(scores, nrows, stride, col, topn, rows);
    if (rc) return rc;
    const size_t cap = topn_cap(nrows, topn);
    VX_REQUIRE(ws && ((uintptr_t)ws & 7) == 0, VX_ERR_ARG, "workspace must be 8-byte aligned device memory");
    VX_REQUIRE((size_t)ws_bytes >= 2 * cap * sizeof(unsigned long long), VX_ERR_ARG,
               "workspace of %lld bytes < the %zu bytes vx_topn_ws_bytes gives", (long long)ws_bytes,
               2 * cap * sizeof(unsigned long long));
    unsigned long long* bufA = (unsigned long long*)ws;
    return topn_run(scores, nrows, stride, col, topn, rows, bufA, bufA + cap, (hipStream_t)stream);
}

extern "C" int vx_mode_planes(int n, int mode, const double* data, const float* vec, const float* prob, double ave,
                              float* out, vx_stream stream) {
    VX_REQUIRE(n > 0 && out && mode >= VX_MODE_VEC && mode <= VX_MODE_PROB, VX_ERR_ARG, "bad arguments");
    const bool needs_vec = mode == VX_MODE_VEC || mode == VX_MODE_PROB;
    VX_REQUIRE(!needs_vec || vec, VX_ERR_ARG, "mode %d needs vec_dev", mode);
    VX_REQUIRE(mode != VX_MODE_PROB || prob, VX_ERR_ARG, "prob mode needs prob_dev");
    VX_REQUIRE(needs_vec || data, VX_ERR_ARG, "mode %d needs data_dev", mode);
    const size_t n3 = (size_t)n * n * n;
    hipLaunchKernelGGL(k_mode_planes, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                       mode, data, vec, prob, ave, out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_laplacian(int n, const double* in, double* out, vx_stream stream) {
    VX_REQUIRE(in && out && n > 0, VX_ERR_ARG, "bad arguments");
    const size_t n3 = (size_t)n * n * n;
    hipLaunchKernelGGL(k_laplacian, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, in, out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

// ---------------------------------------------------------------------------
// Threshold + dmax of mrc_set_vox_size (VESPER_1_prob_new.py:78-85)
// ---------------------------------------------------------------------------

namespace vx {

__device__ __forceinline__ double atomic_max_nonneg(double* addr, double v) {
    // non-negative doubles order like their bit patterns
    unsigned long long* a = (unsigned long long*)addr;
    return __longlong_as_double((long long)atomicMax(a, (unsigned long long)__double_as_longlong(v)));
}

__global__ void __launch_bounds__(256)
k_threshold_dmax(float* __restrict__ d, int sx, int sy, int sz, float th, double* __restrict__ out) {
    __shared__ double s[4];
    const size_t total = (size_t)sx * sy * sz;
    const double cx = sx * 0.5, cy = sy * 0.5, cz = sz * 0.5;
    double best = -1.0;
    double cnt = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        float v = d[i];
        if (v < th) { v = 0.f; d[i] = 0.f; }
        if (v != 0.f) {
            const int z = (int)(i % sz), y = (int)((i / sz) % sy), x = (int)(i / ((size_t)sy * sz));
            const double ex = x - cx, ey = y - cy, ez = z - cz;
            const double r = sqrt((ex * ex + ey * ey) + ez * ez);
            if (r > best) best = r;
            cnt += 1.0;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        best = fmax(best, __shfl_xor(best, off, 64));
        cnt += __shfl_xor(cnt, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __shared__ double sc[4];
    if (lane == 0) { s[wid] = best; sc[wid] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = s[0], c = sc[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { b = fmax(b, s[w]); c += sc[w]; }
        if (b >= 0.0) atomic_max_nonneg(out, b);
        atomicAdd(out + 1, c);
    }
}

__global__ void k_zero2(double* p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0.0;
}

}  // namespace vx

extern "C" int vx_threshold_dmax(float* data, int sx, int sy, int sz, float th, double* out2, vx_stream stream) {
    VX_REQUIRE(data && out2 && sx > 0 && sy > 0 && sz > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    void* scr = nullptr;
    int rc = scratch_get((size_t)kRedBlocks * 8 * sizeof(double) + 64 * sizeof(double), &scr);
    if (rc) return rc;
    double* o = (double*)scr + (size_t)kRedBlocks * 8 + 32;
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, st, o);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_threshold_dmax, dim3(kRedBlocks), dim3(256), 0, st, data, sx, sy, sz, th, o);
    VX_LAUNCH_CHECK();
    double h[2];
    VX_HIP(hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    VX_REQUIRE(h[1] > 0, VX_ERR_ARG, "no voxel above the threshold");
    out2[0] = h[0];
    out2[1] = h[1];
    return VX_OK;
}

// ---------------------------------------------------------------------------
// Reductions over caller-held real correlation maps (the per-row helpers of
// the reference module) and get_score's per-voxel products.
// ---------------------------------------------------------------------------

namespace vx {

constexpr int kMaxMaps = 8;
struct MapSet {
    const float* p[kMaxMaps];
};

// find_best_trans_list (VESPER_1_prob_new.py:365-369): sum_arr = 0; sum_arr =
// sum_arr + arr per channel in float32; amax and the first argmax.
__global__ void __launch_bounds__(256)
k_sum_argmax(MapSet m, int nm, size_t n3, unsigned long long* __restrict__ key) {
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nm; ++c) s = __fadd_rn(s, m.p[c][i]);
        if (better(s, (uint32_t)i, bv, bi)) { bv = s; bi = (uint32_t)i; }
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) atomicMax(key, argmax_key(bv, bi));
}

// find_best_trans_mixed (:1530-1545): Python sum() of each group, float32
// z-scores with float32-rounded constants, alpha mix; no contraction.
__global__ void __launch_bounds__(256)
k_mixed_argmax(MapSet v, int nv, MapSet p, int np, size_t n3, float vave, float vstd, float pave, float pstd,
               float wa, float wb, unsigned long long* __restrict__ key) {
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        float a = 0.f, b = 0.f;
        for (int c = 0; c < nv; ++c) a = __fadd_rn(a, v.p[c][i]);
        for (int c = 0; c < np; ++c) b = __fadd_rn(b, p.p[c][i]);
        const float sv = __fdiv_rn(__fsub_rn(a, vave), vstd);
        const float sp = __fdiv_rn(__fsub_rn(b, pave), pstd);
        const float mx = __fadd_rn(__fmul_rn(wa, sv), __fmul_rn(wb, sp));
        if (better(mx, (uint32_t)i, bv, bi)) { bv = mx; bi = (uint32_t)i; }
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) atomicMax(key, argmax_key(bv, bi));
}

// key (held in the int64 output) -> best value and flat C-order index
__global__ void k_key_out(float* __restrict__ best, int64_t* __restrict__ index) {
    if (threadIdx.x != 0) return;
    const unsigned long long k = (unsigned long long)index[0];
    best[0] = unordered_bits((uint32_t)(k >> 32));
    index[0] = (int64_t)(0xffffffffu - (uint32_t)(k & 0xffffffffu));
}

__global__ void k_zero_u64(int64_t* p) {
    if (threadIdx.x == 0) p[0] = 0;
}

// sco_arr (:1362-1370): at each in-range search position s = t + trans, the
// float32 product sum tv . sv, accumulated as ((0 + x) + y) + z.
template <typename T>
__global__ void __launch_bounds__(256)
k_sco_arr(const float* __restrict__ tvec, const float* __restrict__ svec, int n, int t0, int t1, int t2,
          T* __restrict__ out) {
    const size_t n3 = (size_t)n * n * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % n), y = (int)((i / n) % n), x = (int)(i / ((size_t)n * n));
        const int sx = x + t0, sy = y + t1, sz = z + t2;
        if (sx < 0 || sy < 0 || sz < 0 || sx >= n || sy >= n || sz >= n) continue;
        const size_t j = ((size_t)sx * n + sy) * n + sz;
        float a = 0.f;
        a = __fadd_rn(a, __fmul_rn(tvec[3 * i + 0], svec[3 * j + 0]));
        a = __fadd_rn(a, __fmul_rn(tvec[3 * i + 1], svec[3 * j + 1]));
        a = __fadd_rn(a, __fmul_rn(tvec[3 * i + 2], svec[3 * j + 2]));
        out[j] = (T)a;
    }
}

}  // namespace vx

static int fill_mapset(vx::MapSet* s, const float* const* maps, int nm) {
    VX_REQUIRE(maps != nullptr && nm >= 1 && nm <= vx::kMaxMaps, VX_ERR_ARG, "need 1..%d maps", vx::kMaxMaps);
    for (int c = 0; c < vx::kMaxMaps; ++c) s->p[c] = c < nm ? maps[c] : nullptr;
    for (int c = 0; c < nm; ++c) VX_REQUIRE(maps[c] != nullptr, VX_ERR_ARG, "map %d is NULL", c);
    return VX_OK;
}

extern "C" int vx_sum_argmax(const float* const* maps, int nmaps, int64_t n3, float* best, int64_t* index,
                             vx_stream stream) {
    VX_REQUIRE(best && index && n3 > 0 && n3 < 0xffffffffLL, VX_ERR_ARG, "bad arguments");
    vx::MapSet m;
    int rc = fill_mapset(&m, maps, nmaps);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(vx::k_zero_u64, dim3(1), dim3(64), 0, st, index);
    VX_LAUNCH_CHECK();
    const unsigned g = (unsigned)std::min<int64_t>((n3 + 255) / 256, 2048);
    hipLaunchKernelGGL(vx::k_sum_argmax, dim3(g), dim3(256), 0, st, m, nmaps, (size_t)n3,
                       (unsigned long long*)index);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(vx::k_key_out, dim3(1), dim3(64), 0, st, best, index);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_mixed_argmax(const float* const* vmaps, int nv, const float* const* pmaps, int np, int64_t n3,
                               double alpha, double vstd, double vave, double pstd, double pave, float* best,
                               int64_t* index, vx_stream stream) {
    VX_REQUIRE(best && index && n3 > 0 && n3 < 0xffffffffLL, VX_ERR_ARG, "bad arguments");
    vx::MapSet v, p;
    int rc = fill_mapset(&v, vmaps, nv);
    if (rc) return rc;
    rc = fill_mapset(&p, pmaps, np);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(vx::k_zero_u64, dim3(1), dim3(64), 0, st, index);
    VX_LAUNCH_CHECK();
    const unsigned g = (unsigned)std::min<int64_t>((n3 + 255) / 256, 2048);
    hipLaunchKernelGGL(vx::k_mixed_argmax, dim3(g), dim3(256), 0, st, v, nv, p, np, (size_t)n3, (float)vave,
                       (float)vstd, (float)pave, (float)pstd, (float)(1.0 - alpha), (float)alpha,
                       (unsigned long long*)index);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(vx::k_key_out, dim3(1), dim3(64), 0, st, best, index);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_sco_arr(int n, const float* tvec, const float* svec, const int32_t* trans, int out_f64,
                          void* out, vx_stream stream) {
    VX_REQUIRE(n > 0 && tvec && svec && trans && out, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const size_t n3 = (size_t)n * n * n;
    int t[3];
    for (int a = 0; a < 3; ++a) t[a] = trans[a] > 0.5 * n ? trans[a] - n : trans[a];
    VX_HIP(hipMemsetAsync(out, 0, n3 * (out_f64 ? sizeof(double) : sizeof(float)), st));
    const unsigned g = (unsigned)std::min<size_t>((n3 + 255) / 256, 4096);
    if (out_f64)
        hipLaunchKernelGGL(vx::k_sco_arr<double>, dim3(g), dim3(256), 0, st, tvec, svec, n, t[0], t[1], t[2],
                           (double*)out);
    else
        hipLaunchKernelGGL(vx::k_sco_arr<float>, dim3(g), dim3(256), 0, st, tvec, svec, n, t[0], t[1], t[2],
                           (float*)out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}
