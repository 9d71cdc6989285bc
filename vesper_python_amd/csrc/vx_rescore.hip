// FP64 re-scoring of rotation rows.
//
// The float32 search (vx_score_rows) ranks every row by the maximum of its
// float32 correlation map.  For the rows whose rank matters -- the ones near
// the top-N cutoff and the local-refinement rows -- the score and the
// translation are re-decided in FP64: every translation whose float32 value
// lies within a relative window of the row's float32 maximum is a candidate,
// and each candidate's correlation is summed directly in FP64,
//   corr[t] = N^3 sum_v T[v] Q[(v + t) mod N] = N^3 sum_u Q[u] T[(u - t) mod N],
// over the query's support box, from the caller's target map (its mode
// channels in float64, as the reference's complex128 target spectra hold
// them, VESPER_1_prob_new.py:601-628, :846-860) and the rotated query with
// its vectors rotated in FP64.  The best candidate (largest FP64 value, then
// the smallest C-order index) is the row's FP64 score and translation.  The
// result is the FP64 argmax whenever the float32 map's error stays inside the
// window.
#include <hip/hip_runtime.h>
#include <math.h>
#include <algorithm>

#include "vx_common.hpp"
#include "vx_internal.hpp"
#include "vx_gather.hpp"

namespace vx {

__device__ __forceinline__ double fma_d(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The rotated query at destination voxel (x, y, z), selected as rot_mrc /
// rot_mrc_prob do (float32 positions in the reference's order, the later row
// wins, density and each probability channel by their own non-zero masks),
// with the vector rotated in FP64: new_j = sum_k v_k F[k][j].
__device__ __forceinline__ void gather64(int x, int y, int z, int n, const float* F, bool fma,
                                         const double* __restrict__ data, const float* __restrict__ vec,
                                         const float* __restrict__ prob, double& od, double* ov, double* op) {
    int cx[2], cy[2], cz[2];
    const int nx = dest_candidates(x, n, cx);
    const int ny = dest_candidates(y, n, cy);
    const int nz = dest_candidates(z, n, cz);
    const float cent = 0.5f * (float)n;
    const float fn = (float)n;
    const size_t n3 = (size_t)n * n * n;
    od = 0.0;
    ov[0] = ov[1] = ov[2] = 0.0;
    op[0] = op[1] = op[2] = op[3] = 0.0;
    bool done_v = false;
    bool done_p[4] = {prob == nullptr, prob == nullptr, prob == nullptr, prob == nullptr};
    for (int a = 0; a < nz; ++a)
        for (int b = 0; b < nx; ++b)
            for (int c = 0; c < ny; ++c) {
                const float n0 = (float)cx[b] - cent, n1 = (float)cy[c] - cent, n2 = (float)cz[a] - cent;
                const float o0 = rot_coord(n0, n1, n2, F + 0, cent, fma);
                const float o1 = rot_coord(n0, n1, n2, F + 3, cent, fma);
                const float o2 = rot_coord(n0, n1, n2, F + 6, cent, fma);
                if (!(o0 >= 0.f && o1 >= 0.f && o2 >= 0.f && o0 < fn && o1 < fn && o2 < fn)) continue;
                const size_t si = ((size_t)(int)o0 * n + (int)o1) * n + (int)o2;
                if (!done_v && data[si] != 0.0) {
                    od = data[si];
                    if (vec) {
                        const double s0 = vec[3 * si], s1 = vec[3 * si + 1], s2 = vec[3 * si + 2];
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            ov[j] = fma_d(s2, (double)F[6 + j], fma_d(s1, (double)F[3 + j], s0 * (double)F[j]));
                    }
                    done_v = true;
                }
                for (int k = 0; k < 4; ++k) {
                    if (!done_p[k]) {
                        const float pv = prob[k * n3 + si];
                        if (pv != 0.f) { op[k] = pv; done_p[k] = true; }
                    }
                }
                if (done_v && done_p[0] && done_p[1] && done_p[2] && done_p[3]) return;
            }
}

// Rotated density only (the Laplacian mode's neighbours).
__device__ __forceinline__ double gather_data64(int x, int y, int z, int n, const float* F, bool fma,
                                                const double* __restrict__ data) {
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return 0.0;
    double od, ov[3], op[4];
    gather64(x, y, z, n, F, fma, data, nullptr, nullptr, od, ov, op);
    return od;
}

// laplacian_filter at one voxel of a float64 grid (VESPER_1_prob_new.py:
// 509-525, the reference's addition order; neighbours outside are 0).
template <class GET>
__device__ __forceinline__ double laplacian64(int x, int y, int z, int n, GET get) {
    const double c = get(x, y, z);
    if (!(c > 0)) return 0.0;
    double r = -6.0 * c;
    if (x + 1 < n) r += get(x + 1, y, z);
    if (x - 1 >= 0) r += get(x - 1, y, z);
    if (y + 1 < n) r += get(x, y + 1, z);
    if (y - 1 >= 0) r += get(x, y - 1, z);
    if (z + 1 < n) r += get(x, y, z + 1);
    if (z - 1 >= 0) r += get(x, y, z - 1);
    return r;
}

// The mode transform of a density value (Overlap / CC / PCC,
// VESPER_1_prob_new.py:548-555, :608-615).
__device__ __forceinline__ double mode_value(int mode, double d, double ave) {
    if (!(d > 0)) return 0.0;
    if (mode == VX_MODE_OVERLAP) return 1.0;
    if (mode == VX_MODE_CC) return d;
    return d - ave;
}

// Channels of a group: vector product and prob mode's vector group use the
// 3 vector components, the probability group the 4 probability channels,
// the density modes one channel.
constexpr int kRsSplit = 8;   // workgroups per candidate (partial sums over slabs of the box)

struct RsArgs {
    int n, mode, lo, w, fma, gout, max_cand, r0;
    const double* tdata;
    const float* tvec;
    const float* tprob;
    double ave;
    const double* sdata;
    const float* svec;
    const float* sprob;
    const float* rot;             // [rows][9]
    const int* ncand;             // [rows][gout]
    const int* cand;              // [rows][gout][max_cand] flat C-order translation indices
    double* partial;              // [rows][gout][max_cand][kRsSplit][2]
};

// One workgroup per (candidate slot x split, output group, row of the batch):
// FP64 sums over its share of the query's support box of the vector (or
// density) group s[0] and the probability group s[1].
__global__ void __launch_bounds__(256) k_rs_direct(RsArgs a) {
    const int slot = blockIdx.x / kRsSplit, split = blockIdx.x - slot * kRsSplit;
    const int g = blockIdx.y;
    const int row = a.r0 + blockIdx.z;
    const int nc = min(a.ncand[row * a.gout + g], a.max_cand);
    if (slot >= nc) return;
    const int n = a.n;
    const size_t n3 = (size_t)n * n * n;
    const int idx = a.cand[((size_t)row * a.gout + g) * a.max_cand + slot];
    const int tz = idx % n, ty = (idx / n) % n, tx = idx / (n * n);
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = a.rot[(size_t)row * 9 + k];
    const bool fma = a.fma != 0;
    const bool prob = a.mode == VX_MODE_PROB;
    // prob mode: group 0 needs the vector channels, group 1 the probability
    // channels, the mixed group both
    const bool want_v = !prob || g != 1;
    const bool want_p = prob && g != 0;
    const int w = a.w;
    const long long wb = (long long)w * w * w;
    const long long e0 = wb * split / kRsSplit, e1 = wb * (split + 1) / kRsSplit;
    double s0 = 0.0, s1 = 0.0;
    for (long long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int uz = a.lo + (int)(e % w), uy = a.lo + (int)((e / w) % w), ux = a.lo + (int)(e / ((long long)w * w));
        // target voxel (u - t) mod N
        int vx = ux - tx, vy = uy - ty, vz = uz - tz;
        vx += vx < 0 ? n : 0;
        vy += vy < 0 ? n : 0;
        vz += vz < 0 ? n : 0;
        const size_t it = ((size_t)vx * n + vy) * n + vz;
        if (a.mode == VX_MODE_LAPLACIAN) {
            const double q = laplacian64(ux, uy, uz, n, [&](int x, int y, int z) {
                return gather_data64(x, y, z, n, F, fma, a.sdata);
            });
            if (q != 0.0) {
                const double t = laplacian64(vx, vy, vz, n, [&](int x, int y, int z) {
                    return a.tdata[((size_t)x * n + y) * n + z];
                });
                s0 = fma_d(q, t, s0);
            }
            continue;
        }
        double od, ov[3], op[4];
        gather64(ux, uy, uz, n, F, fma, a.sdata, (a.mode == VX_MODE_VEC || prob) ? a.svec : nullptr,
                 prob ? a.sprob : nullptr, od, ov, op);
        if (a.mode == VX_MODE_VEC || prob) {
            if (want_v && od != 0.0) {
                s0 = fma_d(ov[0], (double)a.tvec[3 * it], s0);
                s0 = fma_d(ov[1], (double)a.tvec[3 * it + 1], s0);
                s0 = fma_d(ov[2], (double)a.tvec[3 * it + 2], s0);
            }
            if (want_p)
                for (int k = 0; k < 4; ++k)
                    if (op[k] != 0.0) s1 = fma_d(op[k], (double)a.tprob[k * n3 + it], s1);
        } else {
            const double q = mode_value(a.mode, od, a.ave);
            if (q != 0.0) s0 = fma_d(q, mode_value(a.mode, a.tdata[it], a.ave), s0);
        }
    }
    // fixed-order block reduction (repeatable sums)
    __shared__ double red[2][4];
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { red[0][wid] = s0; red[1][wid] = s1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) { t0 += red[0][i]; t1 += red[1][i]; }
        double* out = a.partial + ((((size_t)row * a.gout + g) * a.max_cand + slot) * kRsSplit + split) * 2;
        out[0] = t0;
        out[1] = t1;
    }
}

// Candidates of a batch: every translation whose float32 value (the group's
// correlation map, or prob mode's mixed score formed as the float32 search
// forms it) is >= best - window |best| of that row and group.
__global__ void __launch_bounds__(256)
k_rs_candidates(const float* __restrict__ maps, int G, size_t n3, int gout, int r0, const float* __restrict__ best32,
                double window, float vave, float vstd, float pave, float pstd, float wa, float wb, int max_cand,
                int* __restrict__ ncand, int* __restrict__ cand) {
    const int b = blockIdx.y;
    const int row = r0 + b;
    for (int g = 0; g < gout; ++g) {
        const double bv = (double)best32[(size_t)row * gout + g];
        const double thr = bv - window * fabs(bv);
        const float* m0 = maps + (size_t)b * G * n3;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
            float v;
            if (g < G) v = m0[(size_t)g * n3 + i];
            else {
                const float sv = __fdiv_rn(__fsub_rn(m0[i], vave), vstd);
                const float sp = __fdiv_rn(__fsub_rn(m0[n3 + i], pave), pstd);
                v = __fadd_rn(__fmul_rn(wa, sv), __fmul_rn(wb, sp));
            }
            if ((double)v >= thr) {
                const int slot = atomicAdd(&ncand[(size_t)row * gout + g], 1);
                if (slot < max_cand) cand[((size_t)row * gout + g) * max_cand + slot] = (int)i;
            }
        }
    }
}

// Per (row, group): the best candidate by (FP64 value desc, flat index asc).
__global__ void k_rs_select(int nrows, int gout, int n, int max_cand, int prob, const int* __restrict__ ncand,
                            const int* __restrict__ cand, const double* __restrict__ partial, double alpha,
                            double vstd, double vave, double pstd, double pave, double* __restrict__ score,
                            int32_t* __restrict__ trans) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows * gout) return;
    const int g = i % gout;
    const int nc = min(ncand[i], max_cand);
    const double n3 = (double)n * n * n;
    double best = -INFINITY;
    int bidx = -1;
    for (int s = 0; s < nc; ++s) {
        double t0 = 0.0, t1 = 0.0;
        const double* pp = partial + (((size_t)i * max_cand + s) * kRsSplit) * 2;
        for (int k = 0; k < kRsSplit; ++k) { t0 += pp[2 * k]; t1 += pp[2 * k + 1]; }
        double v;
        if (!prob || g == 0) v = n3 * t0;
        else if (g == 1) v = n3 * t1;
        else v = (1.0 - alpha) * (n3 * t0 - vave) / vstd + alpha * (n3 * t1 - pave) / pstd;
        const int idx = cand[(size_t)i * max_cand + s];
        if (v > best || (v == best && idx < bidx)) { best = v; bidx = idx; }
    }
    score[i] = bidx >= 0 ? best : NAN;
    trans[3 * i + 0] = bidx >= 0 ? bidx / (n * n) : -1;
    trans[3 * i + 1] = bidx >= 0 ? (bidx / n) % n : -1;
    trans[3 * i + 2] = bidx >= 0 ? bidx % n : -1;
}

int rs_launch_candidates(const float* maps, int G, int n, int nb, int gout, int r0, const float* best32,
                         double window, const RsStats& st, int max_cand, int* ncand, int* cand, hipStream_t s) {
    const size_t n3 = (size_t)n * n * n;
    const unsigned gx = (unsigned)std::min<size_t>((n3 + 255) / 256, 1024);
    hipLaunchKernelGGL(k_rs_candidates, dim3(gx, nb), dim3(256), 0, s, maps, G, n3, gout, r0, best32, window,
                       (float)st.vave, (float)st.vstd, (float)st.pave, (float)st.pstd, (float)(1.0 - st.alpha),
                       (float)st.alpha, max_cand, ncand, cand);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

int rs_launch_direct(const RsInputs& in, int n, int lo, int w, int fma, int gout, int max_cand, int r0, int nb,
                     const int* ncand, const int* cand, double* partial, hipStream_t s) {
    RsArgs a;
    a.n = n; a.mode = in.mode; a.lo = lo; a.w = w; a.fma = fma; a.gout = gout; a.max_cand = max_cand; a.r0 = r0;
    a.tdata = in.tdata; a.tvec = in.tvec; a.tprob = in.tprob; a.ave = in.ave;
    a.sdata = in.sdata; a.svec = in.svec; a.sprob = in.sprob; a.rot = in.rot;
    a.ncand = ncand; a.cand = cand; a.partial = partial;
    hipLaunchKernelGGL(k_rs_direct, dim3(max_cand * kRsSplit, gout, nb), dim3(256), 0, s, a);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

int rs_launch_select(int nrows, int gout, int n, int max_cand, int prob, const int* ncand, const int* cand,
                     const double* partial, const RsStats& st, double* score, int32_t* trans, hipStream_t s) {
    const int m = nrows * gout;
    hipLaunchKernelGGL(k_rs_select, dim3((m + 127) / 128), dim3(128), 0, s, nrows, gout, n, max_cand, prob, ncand,
                       cand, partial, st.alpha, st.vstd, st.vave, st.pstd, st.pave, score, trans);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

size_t rs_partial_doubles(int nrows, int gout, int max_cand) {
    return (size_t)nrows * gout * max_cand * kRsSplit * 2;
}

}  // namespace vx
