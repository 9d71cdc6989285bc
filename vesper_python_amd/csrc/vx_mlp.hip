// Learned rotation proposal on MI355X: a 4-layer MLP (width 256, bf16
// weights and activations, fp32 accumulation on MFMA) that predicts a
// rotation's DOT score from its quaternion and low-order rotational moments of
// the two vector fields, trained by these kernels on score landscapes the
// exhaustive search produces.  No reference counterpart (the reference
// scores every rotation, VESPER_1_prob_new.py:644-655).
//
//   vx_gemm_bf16        C = op(A) op(B) on v_mfma_f32_16x16x16_bf16, fused
//                       epilogues (bias + ReLU -> bf16, ReLU mask -> bf16,
//                       fp32 store, fp32 atomic accumulate)
//   vx_mlp_head_*       the 256 -> 1 output layer, forward and backward
//   vx_colsum_bf16      bias gradients
//   vx_adam             optimiser step on fp32 master weights + bf16 copy
//   vx_mlp_descriptors  32 rotational moments of one resampled map
//   vx_mlp_features     per rotation row: quaternion, target moments,
//                       moments of the rotated query -> bf16 input row
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <math.h>

#include "vx_common.hpp"
#include "vx_internal.hpp"

namespace vx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// One 64x64 block tile per 256-thread workgroup; each wave owns a 32x32
// quarter as 2x2 MFMA 16x16 tiles.  Operand layout of the 16x16x16 bf16
// MFMA: lane l holds A[l % 16][4 (l / 16) + v] and B[4 (l / 16) + v][l % 16]
// (v = 0..3); the accumulator lane l holds D[4 (l / 16) + v][l % 16].
// K must be a multiple of 16; rows/columns past M/N read as zero and are not written.
template <int TA, int TB>
__global__ void __launch_bounds__(256)
k_gemm_bf16(int M, int N, int K, const unsigned short* __restrict__ A, int lda,
            const unsigned short* __restrict__ B, int ldb, void* __restrict__ C, int ldc, int epi,
            const float* __restrict__ bias, const unsigned short* __restrict__ aux, int ldaux) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int c0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int li = lane & 15, lk = (lane >> 4) * 4;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int row = r0 + 16 * i + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kk = k0 + lk + v;
                a[i][v] = row < M ? (short)(TA ? A[(size_t)kk * lda + row] : A[(size_t)row * lda + kk]) : (short)0;
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = c0 + 16 * j + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kk = k0 + lk + v;
                b[j][v] = col < N ? (short)(TB ? B[(size_t)col * ldb + kk] : B[(size_t)kk * ldb + col]) : (short)0;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r0 + 16 * i + (lane >> 4) * 4 + v;
                const int col = c0 + 16 * j + li;
                if (row >= M || col >= N) continue;
                float x = acc[i][j][v];
                const size_t o = (size_t)row * ldc + col;
                if (epi == 0) {                       // fp32 store
                    ((float*)C)[o] = x;
                } else if (epi == 1) {                // bias + ReLU -> bf16
                    x += bias[col];
                    ((unsigned short*)C)[o] = f2bf(x > 0.f ? x : 0.f);
                } else if (epi == 2) {                // gradient through ReLU of aux -> bf16
                    const float m = bf2f(aux[(size_t)row * ldaux + col]);
                    ((unsigned short*)C)[o] = f2bf(m > 0.f ? x : 0.f);
                } else {                              // fp32 atomic accumulate
                    atomicAdd(&((float*)C)[o], x);
                }
            }
}

// Output layer 256 -> 1: pred[m] = sum_k X[m][k] w[k] + b
__global__ void k_head_fwd(int M, int K, const unsigned short* __restrict__ X, int ldx, const float* __restrict__ w,
                           const float* __restrict__ b, float* __restrict__ pred) {
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (m >= M) return;
    float s = 0.f;
    for (int k = lane; k < K; k += 64) s += bf2f(X[(size_t)m * ldx + k]) * w[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) pred[m] = s + b[0];
}

// Backward of the output layer with the mean-squared loss, d = 2 (pred - y) / M:
// dX[m][k] = d w[k] * (X[m][k] > 0)  (through the ReLU that produced X),
// one wave per row.
__global__ void k_head_bwd_rows(int M, int K, const unsigned short* __restrict__ X, int ldx,
                                const float* __restrict__ w, const float* __restrict__ pred,
                                const float* __restrict__ y, unsigned short* __restrict__ dX) {
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (m >= M) return;
    const float d = 2.f * (pred[m] - y[m]) / (float)M;
    for (int k = lane; k < K; k += 64) {
        const float xv = bf2f(X[(size_t)m * ldx + k]);
        dX[(size_t)m * ldx + k] = f2bf(xv > 0.f ? d * w[k] : 0.f);
    }
}

// Row sums in a fixed order, so that repeated runs give identical results:
// a 1024-thread workgroup holds 64 columns x 16 row phases; phase p sums rows
// p, p + 16, ... in order, then phase 0 adds the 16 partials in order.
constexpr int kSumPhases = 16;

// dw[k] += sum_m d_m X[m][k];  db += sum_m d_m;  loss += sum_m e_m^2 / M
// (e = pred - y, d = 2 e / M); db and loss come from workgroup 0.
__global__ void __launch_bounds__(1024)
k_head_bwd_sums(int M, int K, const unsigned short* __restrict__ X, int ldx, const float* __restrict__ pred,
                const float* __restrict__ y, float* __restrict__ dw, float* __restrict__ db,
                float* __restrict__ loss) {
    __shared__ float red[3][kSumPhases][64];
    const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + c;
    float s = 0.f, sd = 0.f, se = 0.f;
#pragma unroll 4
    for (int m = ph; m < M; m += kSumPhases) {
        const float e = pred[m] - y[m];
        const float d = 2.f * e / (float)M;
        if (k < K) s += d * bf2f(X[(size_t)m * ldx + k]);
        sd += d;
        se += e * e / (float)M;
    }
    red[0][ph][c] = s;
    red[1][ph][c] = sd;
    red[2][ph][c] = se;
    __syncthreads();
    if (ph != 0) return;
    float a = 0.f, b = 0.f, l = 0.f;
    for (int p = 0; p < kSumPhases; ++p) {
        a += red[0][p][c];
        b += red[1][p][c];
        l += red[2][p][c];
    }
    if (k < K) dw[k] += a;
    if (blockIdx.x == 0 && c == 0) {
        db[0] += b;
        loss[0] += l;
    }
}

// out[col] += sum_m G[m][col]  (bias gradients), fixed summation order.
__global__ void __launch_bounds__(1024)
k_colsum_bf16(int M, int N, const unsigned short* __restrict__ G, int ldg, float* __restrict__ out) {
    __shared__ float red[kSumPhases][64];
    const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + c;
    float s = 0.f;
    if (col < N) {
#pragma unroll 8
        for (int m = ph; m < M; m += kSumPhases) s += bf2f(G[(size_t)m * ldg + col]);
    }
    red[ph][c] = s;
    __syncthreads();
    if (ph != 0 || col >= N) return;
    float a = 0.f;
    for (int p = 0; p < kSumPhases; ++p) a += red[p][c];
    out[col] += a;
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, unsigned short* __restrict__ pb, float lr, float b1, float b2,
                       float eps, float c1, float c2, float wd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    if (pb) pb[i] = f2bf(v);
}

__global__ void k_cast_bf16(int n, const float* __restrict__ in, unsigned short* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// 32 moments of a resampled map about the rotation centre n/2, positions in
// units of n: w = data > 0, r = (index - n/2) / n, v = vec.
//  sums: [0] sum w, [1..3] sum w r, [4..9] sum w r r (xx yy zz xy xz yz),
//        [10] count, [11..13] sum v, [14..22] sum r v^T (row-major),
//        [23..28] sum v v (xx yy zz xy xz yz)
__global__ void __launch_bounds__(256)
k_moment_sums(int n, const double* __restrict__ data, const float* __restrict__ vec, double* __restrict__ out) {
    double s[29];
#pragma unroll
    for (int k = 0; k < 29; ++k) s[k] = 0.0;
    const size_t n3 = (size_t)n * n * n;
    const double c = 0.5 * n, inv = 1.0 / n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const double d = data[i];
        if (!(d > 0.0)) continue;
        const int z = (int)(i % n), y = (int)((i / n) % n), x = (int)(i / ((size_t)n * n));
        const double r[3] = {(x - c) * inv, (y - c) * inv, (z - c) * inv};
        const double v[3] = {vec[3 * i], vec[3 * i + 1], vec[3 * i + 2]};
        s[0] += d;
        for (int a = 0; a < 3; ++a) s[1 + a] += d * r[a];
        s[4] += d * r[0] * r[0]; s[5] += d * r[1] * r[1]; s[6] += d * r[2] * r[2];
        s[7] += d * r[0] * r[1]; s[8] += d * r[0] * r[2]; s[9] += d * r[1] * r[2];
        s[10] += 1.0;
        for (int a = 0; a < 3; ++a) s[11 + a] += v[a];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s[14 + 3 * a + b] += r[a] * v[b];
        s[23] += v[0] * v[0]; s[24] += v[1] * v[1]; s[25] += v[2] * v[2];
        s[26] += v[0] * v[1]; s[27] += v[0] * v[2]; s[28] += v[1] * v[2];
    }
    // per-workgroup partials in a fixed order (out[blockIdx.x * 32 + k]);
    // k_moment_finish adds the workgroups' partials in order
    __shared__ double wsum[4][29];
#pragma unroll
    for (int k = 0; k < 29; ++k) {
        double t = s[k];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 29)
        out[(size_t)blockIdx.x * 32 + threadIdx.x] =
            ((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) + wsum[3][threadIdx.x];
}

// sums -> 32 descriptors: centroid (3), second moments (6), mean vector (3),
// position-vector moments (9), vector second moments (6), invariants (5:
// trace of second moments, |centroid|, |mean vector|, trace of r v^T,
// occupancy)
__global__ void k_moment_finish(int n, int nblk, const double* __restrict__ part, float* __restrict__ d) {
    __shared__ double s[29];
    if (threadIdx.x < 29) {
        double t = 0.0;
        for (int b = 0; b < nblk; ++b) t += part[(size_t)b * 32 + threadIdx.x];
        s[threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double w = s[0] > 0 ? s[0] : 1.0, c = s[10] > 0 ? s[10] : 1.0;
    for (int a = 0; a < 3; ++a) d[a] = (float)(s[1 + a] / w);
    for (int a = 0; a < 6; ++a) d[3 + a] = (float)(s[4 + a] / w);
    for (int a = 0; a < 3; ++a) d[9 + a] = (float)(s[11 + a] / c);
    for (int a = 0; a < 9; ++a) d[12 + a] = (float)(s[14 + a] / c);
    for (int a = 0; a < 6; ++a) d[21 + a] = (float)(s[23 + a] / c);
    d[27] = d[3] + d[4] + d[5];
    d[28] = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[29] = sqrtf(d[9] * d[9] + d[10] * d[10] + d[11] * d[11]);
    d[30] = d[12] + d[16] + d[20];
    d[31] = (float)(s[10] / ((double)n * n * n));
}

__device__ __forceinline__ void sym_to_mat(const float* s, float m[3][3]) {
    m[0][0] = s[0]; m[1][1] = s[1]; m[2][2] = s[2];
    m[0][1] = m[1][0] = s[3]; m[0][2] = m[2][0] = s[4]; m[1][2] = m[2][1] = s[5];
}

// R M R^T
__device__ __forceinline__ void rot2(const float R[3][3], const float M[3][3], float out[3][3]) {
    float t[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) t[a][b] = R[a][0] * M[0][b] + R[a][1] * M[1][b] + R[a][2] * M[2][b];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) out[a][b] = t[a][0] * R[b][0] + t[a][1] * R[b][1] + t[a][2] * R[b][2];
}

// Feature row: [quaternion (4), target moments (32), moments of the query
// rotated by R = F^T (32), zero padding] in bf16, F = float32(flip(M)) row.
__global__ void k_features(int M, int ld, const float* __restrict__ rot, const float* __restrict__ quat,
                           const float* __restrict__ td, const float* __restrict__ qd,
                           unsigned short* __restrict__ X) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float R[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) R[a][b] = rot[(size_t)m * 9 + 3 * b + a];   // R = F^T
    float f[68];
    for (int a = 0; a < 4; ++a) f[a] = quat[(size_t)m * 4 + a];
    for (int a = 0; a < 32; ++a) f[4 + a] = td[a];
    float* g = f + 36;
    // vectors rotate by R
    for (int a = 0; a < 3; ++a) g[a] = R[a][0] * qd[0] + R[a][1] * qd[1] + R[a][2] * qd[2];
    for (int a = 0; a < 3; ++a) g[9 + a] = R[a][0] * qd[9] + R[a][1] * qd[10] + R[a][2] * qd[11];
    float S[3][3], T[3][3];
    sym_to_mat(qd + 3, S);
    rot2(R, S, T);
    g[3] = T[0][0]; g[4] = T[1][1]; g[5] = T[2][2]; g[6] = T[0][1]; g[7] = T[0][2]; g[8] = T[1][2];
    float P[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] = qd[12 + 3 * a + b];
    rot2(R, P, T);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) g[12 + 3 * a + b] = T[a][b];
    sym_to_mat(qd + 21, S);
    rot2(R, S, T);
    g[21] = T[0][0]; g[22] = T[1][1]; g[23] = T[2][2]; g[24] = T[0][1]; g[25] = T[0][2]; g[26] = T[1][2];
    for (int a = 27; a < 32; ++a) g[a] = qd[a];            // invariants
    unsigned short* row = X + (size_t)m * ld;
    for (int a = 0; a < ld; ++a) row[a] = a < 68 ? f2bf(f[a]) : (unsigned short)0;
}

}  // namespace vx

using namespace vx;

extern "C" int vx_gemm_bf16(int M, int N, int K, const void* A, int lda, int transA, const void* B, int ldb,
                            int transB, void* C, int ldc, int epilogue, const float* bias, const void* aux,
                            int ldaux, vx_stream stream) {
    VX_REQUIRE(A && B && C && M > 0 && N > 0 && K > 0, VX_ERR_ARG, "bad gemm arguments");
    VX_REQUIRE(K % 16 == 0, VX_ERR_ARG, "gemm K must be a multiple of 16");
    VX_REQUIRE(epilogue >= 0 && epilogue <= 3, VX_ERR_ARG, "bad epilogue");
    VX_REQUIRE(epilogue != 1 || bias, VX_ERR_ARG, "bias required");
    VX_REQUIRE(epilogue != 2 || aux, VX_ERR_ARG, "aux required");
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    const unsigned short* a = (const unsigned short*)A;
    const unsigned short* b = (const unsigned short*)B;
    const unsigned short* x = (const unsigned short*)aux;
    hipStream_t st = (hipStream_t)stream;
    if (!transA && !transB) hipLaunchKernelGGL((k_gemm_bf16<0, 0>), grid, dim3(256), 0, st, M, N, K, a, lda, b, ldb, C, ldc, epilogue, bias, x, ldaux);
    else if (!transA)       hipLaunchKernelGGL((k_gemm_bf16<0, 1>), grid, dim3(256), 0, st, M, N, K, a, lda, b, ldb, C, ldc, epilogue, bias, x, ldaux);
    else if (!transB)       hipLaunchKernelGGL((k_gemm_bf16<1, 0>), grid, dim3(256), 0, st, M, N, K, a, lda, b, ldb, C, ldc, epilogue, bias, x, ldaux);
    else                    hipLaunchKernelGGL((k_gemm_bf16<1, 1>), grid, dim3(256), 0, st, M, N, K, a, lda, b, ldb, C, ldc, epilogue, bias, x, ldaux);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_mlp_head_fwd(int M, int K, const void* X, int ldx, const float* w, const float* b, float* pred,
                               vx_stream stream) {
    VX_REQUIRE(X && w && b && pred && M > 0 && K > 0, VX_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_head_fwd, dim3((M + 3) / 4), dim3(256), 0, (hipStream_t)stream, M, K,
                       (const unsigned short*)X, ldx, w, b, pred);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_mlp_head_bwd(int M, int K, const void* X, int ldx, const float* w, const float* pred,
                               const float* y, float* dw, float* db, void* dX, float* loss, vx_stream stream) {
    VX_REQUIRE(X && w && pred && y && dw && db && dX && loss && M > 0, VX_ERR_ARG, "bad arguments");
    VX_REQUIRE(K > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_head_bwd_rows, dim3((M + 3) / 4), dim3(256), 0, st, M, K, (const unsigned short*)X, ldx,
                       w, pred, y, (unsigned short*)dX);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_head_bwd_sums, dim3((K + 63) / 64), dim3(1024), 0, st, M, K, (const unsigned short*)X,
                       ldx, pred, y, dw, db, loss);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_colsum_bf16(int M, int N, const void* G, int ldg, float* out, vx_stream stream) {
    VX_REQUIRE(G && out && M > 0 && N > 0, VX_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_colsum_bf16, dim3((N + 63) / 64), dim3(1024), 0, (hipStream_t)stream, M, N,
                       (const unsigned short*)G, ldg, out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_adam(int n, float* p, const float* g, float* m1, float* m2, void* p_bf16, float lr, float b1,
                       float b2, float eps, int step, float wd, vx_stream stream) {
    VX_REQUIRE(p && g && m1 && m2 && n > 0 && step > 0, VX_ERR_ARG, "bad arguments");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, p, g, m1, m2,
                       (unsigned short*)p_bf16, lr, b1, b2, eps, c1, c2, wd);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_cast_bf16(int n, const float* in, void* out, vx_stream stream) {
    VX_REQUIRE(in && out && n > 0, VX_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_cast_bf16, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, in,
                       (unsigned short*)out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

extern "C" int vx_mlp_descriptors(int n, const double* data, const float* vec, float* out32, vx_stream stream) {
    VX_REQUIRE(data && vec && out32 && n > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    constexpr int nblk = 256;   // every workgroup writes its 29 partials: no clearing needed
    double* sums = nullptr;
    VX_HIP(hipMallocAsync((void**)&sums, nblk * 32 * sizeof(double), st));
    hipLaunchKernelGGL(k_moment_sums, dim3(nblk), dim3(256), 0, st, n, data, vec, sums);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_moment_finish, dim3(1), dim3(64), 0, st, n, nblk, sums, out32);
    VX_LAUNCH_CHECK();
    VX_HIP(hipFreeAsync(sums, st));
    return VX_OK;
}

extern "C" int vx_mlp_features(int M, int ld, const float* rot, const float* quat, const float* tdesc,
                               const float* qdesc, void* X, vx_stream stream) {
    VX_REQUIRE(rot && quat && tdesc && qdesc && X && M > 0 && ld >= 68, VX_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_features, dim3((M + 127) / 128), dim3(128), 0, (hipStream_t)stream, M, ld, rot, quat,
                       tdesc, qdesc, (unsigned short*)X);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

// ---------------------------------------------------------------------------
// Top fraction selection: rows whose prediction is at or above the value
// that ranks k-th (through a 4096-bin histogram between min and max, so a
// few more than k rows can pass at the threshold bin).
// ---------------------------------------------------------------------------

namespace vx {

__global__ void k_minmax(int M, const float* __restrict__ v, unsigned* __restrict__ mm) {
    unsigned lo = 0xffffffffu, hi = 0u;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const unsigned o = ordered_bits(v[i]);
        lo = min(lo, o);
        hi = max(hi, o);
    }
    atomicMin(&mm[0], lo);
    atomicMax(&mm[1], hi);
}

__global__ void k_hist(int M, const float* __restrict__ v, const unsigned* __restrict__ mm, unsigned* __restrict__ h) {
    const float lo = unordered_bits(mm[0]), hi = unordered_bits(mm[1]);
    const float scale = hi > lo ? 4096.f / (hi - lo) : 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        int b = (int)((v[i] - lo) * scale);
        b = b < 0 ? 0 : (b > 4095 ? 4095 : b);
        atomicAdd(&h[b], 1u);
    }
}

__global__ void k_init_mm(unsigned* mm) {
    if (threadIdx.x == 0) { mm[0] = 0xffffffffu; mm[1] = 0u; mm[2] = 0u; }
}

}  // namespace vx

extern "C" int vx_select_top(int M, const float* pred, int k, int* rows, int* nsel_host, vx_stream stream) {
    VX_REQUIRE(pred && rows && nsel_host && M > 0 && k > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    unsigned* buf = nullptr;   // [0..2] min, max, count; [4 .. 4100) histogram
    VX_HIP(hipMallocAsync((void**)&buf, (4 + 4096) * sizeof(unsigned), st));
    VX_HIP(hipMemsetAsync(buf, 0, (4 + 4096) * sizeof(unsigned), st));
    hipLaunchKernelGGL(k_init_mm, dim3(1), dim3(64), 0, st, buf);
    hipLaunchKernelGGL(k_minmax, dim3(256), dim3(256), 0, st, M, pred, buf);
    hipLaunchKernelGGL(k_hist, dim3(256), dim3(256), 0, st, M, pred, buf, buf + 4);
    VX_LAUNCH_CHECK();
    unsigned h[4096];
    VX_HIP(hipMemcpyAsync(h, buf + 4, sizeof(h), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    int bin = 4095;
    long acc = 0;
    for (; bin > 0; --bin) {
        acc += h[bin];
        if (acc >= k) break;
    }
    // rows in ascending order (ordered compaction; the histogram buffer is
    // free again and holds the compaction's block counts and total)
    int rc = compact_top_bins(M, pred, buf, bin, rows, buf + 4, st);
    if (rc != VX_OK) return rc;
    unsigned cnt = 0;
    VX_HIP(hipMemcpyAsync(&cnt, buf + 4 + 256, sizeof(cnt), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    VX_HIP(hipFreeAsync(buf, st));
    *nsel_host = (int)cnt;
    return VX_OK;
}
