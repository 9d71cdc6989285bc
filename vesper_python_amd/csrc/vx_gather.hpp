// Index arithmetic of the rotate-and-gather (rot_mrc, VESPER_1_prob_new.py:
// 279-354) shared by the float32 gathers and the FP64 re-scoring.
#pragma once
#include <hip/hip_runtime.h>

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

}  // namespace vx
