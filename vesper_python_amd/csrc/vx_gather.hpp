// Index arithmetic of the rotate-and-gather (rot_mrc, VESPER_1_prob_new.py:
// 279-354) shared by the float32 gather and the FP64 re-scoring.
#pragma once
#include <hip/hip_runtime.h>

#include "vx_internal.hpp"

namespace vx {

// Source rows i (per axis) that land on destination index d after
// int(int32(i - cent) + cent) (VESPER_1_prob_new.py:329, :348), latest first.
// Even n: the identity.  Odd n (h = (n-1)/2): d < h <- d; d == h <- {h+1, h};
// h < d < n-1 <- d+1; d == n-1 <- nothing.
__device__ __forceinline__ int dest_candidates(int d, int n, int* c) {
    if ((n & 1) == 0) { c[0] = d; return 1; }
    const int h = (n - 1) >> 1;
    if (d < h) { c[0] = d; return 1; }
    if (d == h) { c[0] = h + 1; c[1] = h; return 2; }
    if (d < n - 1) { c[0] = d + 1; return 1; }
    return 0;
}

// old_j = sum_k n_k F[j][k] in the K=3 sgemm order, then + cent (float32).
__device__ __forceinline__ float rot_coord(float n0, float n1, float n2, const float* Fj, float cent, bool fma) {
    float t = __fmul_rn(n0, Fj[0]);
    if (fma) {
        t = __builtin_fmaf(n1, Fj[1], t);
        t = __builtin_fmaf(n2, Fj[2], t);
    } else {
        t = __fadd_rn(t, __fmul_rn(n1, Fj[1]));
        t = __fadd_rn(t, __fmul_rn(n2, Fj[2]));
    }
    return __fadd_rn(t, cent);
}

// new_vec_j = sum_k v_k F[k][j]   (VESPER_1_prob_new.py:341)
__device__ __forceinline__ float rot_vec(float v0, float v1, float v2, const float* F, int j, bool fma) {
    float t = __fmul_rn(v0, F[j]);
    if (fma) {
        t = __builtin_fmaf(v1, F[3 + j], t);
        t = __builtin_fmaf(v2, F[6 + j], t);
    } else {
        t = __fadd_rn(t, __fmul_rn(v1, F[3 + j]));
        t = __fadd_rn(t, __fmul_rn(v2, F[6 + j]));
    }
    return t;
}

// Gather the value(s) landing on destination (x, y, z) for rotation F.
//   KIND_VEC:     svec {vx, vy, vz, data != 0}        -> out[0..2]
//   KIND_SCALAR:  sscal {channel value, data != 0}    -> out[0]
//   KIND_PROB:    svec as above + sprob {p1..p4} each masked by its own != 0
//                                                     -> out[0..6]
//   KIND_DATA64:  sdat (float64, masked by != 0)      -> outd
// svec / sscal / sprob hold only the support box [lo, lo + w)^3 of the source
// (x-major, edge w): every voxel with non-zero density or probability lies
// inside it, so a rotated position outside reads nothing, as a zero voxel
// would.  sdat holds the whole grid.
template <int KIND>
__device__ __forceinline__ void gather_voxel(int x, int y, int z, int n, int lo, int w, const float* F, bool fma,
                                             const float4* __restrict__ svec,
                                             const float2* __restrict__ sscal,
                                             const float4* __restrict__ sprob,
                                             const double* __restrict__ sdat,
                                             float* out, double* outd) {
    constexpr int NOUT = (KIND == KIND_VEC) ? 3 : (KIND == KIND_PROB ? 7 : 1);
#pragma unroll
    for (int c = 0; c < NOUT; ++c) out[c] = 0.f;
    *outd = 0.0;
    int cx[2], cy[2], cz[2];
    const int nx = dest_candidates(x, n, cx);
    const int ny = dest_candidates(y, n, cy);
    const int nz = dest_candidates(z, n, cz);
    const float cent = 0.5f * (float)n;
    const float fn = (float)n;
    bool done_v = false;
    bool done_p[4] = {false, false, false, false};
    for (int a = 0; a < nz; ++a)
    for (int b = 0; b < nx; ++b)
    for (int c = 0; c < ny; ++c) {
        const float n0 = (float)cx[b] - cent, n1 = (float)cy[c] - cent, n2 = (float)cz[a] - cent;
        const float o0 = rot_coord(n0, n1, n2, F + 0, cent, fma);
        const float o1 = rot_coord(n0, n1, n2, F + 3, cent, fma);
        const float o2 = rot_coord(n0, n1, n2, F + 6, cent, fma);
        if (!(o0 >= 0.f && o1 >= 0.f && o2 >= 0.f && o0 < fn && o1 < fn && o2 < fn)) continue;
        size_t si;
        if (KIND == KIND_DATA64) {
            si = ((size_t)(int)o0 * n + (int)o1) * n + (int)o2;
        } else {
            const int b0 = (int)o0 - lo, b1 = (int)o1 - lo, b2 = (int)o2 - lo;
            if ((unsigned)b0 >= (unsigned)w || (unsigned)b1 >= (unsigned)w || (unsigned)b2 >= (unsigned)w) continue;
            si = ((size_t)b0 * w + b1) * w + b2;
        }
        if (KIND == KIND_VEC || KIND == KIND_PROB) {
            if (!done_v) {
                const float4 s = svec[si];
                if (s.w != 0.f) {
                    out[0] = rot_vec(s.x, s.y, s.z, F, 0, fma);
                    out[1] = rot_vec(s.x, s.y, s.z, F, 1, fma);
                    out[2] = rot_vec(s.x, s.y, s.z, F, 2, fma);
                    done_v = true;
                }
            }
            if (KIND == KIND_PROB) {
                const float4 p = sprob[si];
                const float pv[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (!done_p[k] && pv[k] != 0.f) { out[3 + k] = pv[k]; done_p[k] = true; }
                if (done_v && done_p[0] && done_p[1] && done_p[2] && done_p[3]) return;
            } else if (done_v) {
                return;
            }
        } else if (KIND == KIND_SCALAR) {
            const float2 s = sscal[si];
            if (s.y != 0.f) { out[0] = s.x; return; }
        } else {
            const double s = sdat[si];
            if (s != 0.0) { *outd = s; return; }
        }
    }
}

}  // namespace vx
