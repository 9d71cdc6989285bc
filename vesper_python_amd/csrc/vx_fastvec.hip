// fastVEC on MI355X: the Gaussian-weighted density and unit centroid vector
// of every resampled voxel (doFastVEC/calc, VESPER_1_prob_new.py:180-257,
// :121-142).
//
// The reference sums src * exp(-1.5 d^2 fsiv) over an axis-aligned window per
// destination voxel.  The window is a product of per-axis intervals and
// exp(-1.5 (rx^2 + ry^2 + rz^2) fsiv) = wx * wy * wz, so the field is an
// exactly separable banded 3-mode contraction.  It is computed in float64 as
// three passes (z, then y, then x), each a banded product against per-axis
// weight tables, instead of the reference's (4 g / w + 1)^3 exp evaluations
// per voxel:
//   A0[xp][yp][dz] = sum_zp wz[dz][zp] s[xp][yp][zp]       A1: same * zp
//   B0[xp][dy][dz] = sum_yp wy[dy][yp] A0                  B1: * yp   B2: from A1
//   D  = sum_xp wx B0,  Mx = sum wx xp B0,  My = sum wx B1,  Mz = sum wx B2
// then data = D and vec = (M / D - pos) / |M / D - pos| where the reference
// computes the voxel (in-bounds and src[int(pos)] != 0).
#include <hip/hip_runtime.h>
#include <math.h>

#include "vx_common.hpp"

namespace vx {

struct AxisGeom {
    double* pos;    // [3][n]
    int* stp;       // [3][n]
    int* len;       // [3][n] window length (0 if empty)
    int* inb;       // [3][n] 0 <= pos < dim
    int* ipos;      // [3][n] int(pos)
    double* w;      // [3][W][n]: tap j of destination voxel d at w[(a W + j) n + d], so that
                    // neighbouring destinations (the lanes of k_fv_z) read neighbouring weights
};

// Per-axis geometry (VESPER_1_prob_new.py:196, :214-232) and weights.
__global__ void k_fv_axis(AxisGeom g, int n, int W, int sx, int sy, int sz, double dest_width,
                          double o0, double o1, double o2, double s0, double s1, double s2,
                          double src_width, double fmaxd, double fsiv) {
    const int a = blockIdx.y;
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const double od = a == 0 ? o0 : (a == 1 ? o1 : o2);
    const double os = a == 0 ? s0 : (a == 1 ? s1 : s2);
    const int dim = a == 0 ? sx : (a == 1 ? sy : sz);
    // ((xyz * dest_width + dest_orig) - src_orig) / src_width, rounded per op
    const double pos = __ddiv_rn(__dsub_rn(__dadd_rn(__dmul_rn((double)d, dest_width), od), os), src_width);
    int stp = (int)(__dsub_rn(pos, fmaxd));
    if (stp < 0) stp = 0;
    int endp = (int)(__dadd_rn(__dadd_rn(pos, fmaxd), 1.0));
    if (endp >= dim) endp = dim;
    int len = endp - stp;
    if (len < 0) len = 0;
    if (len > W) len = W;
    const int i = a * n + d;
    g.pos[i] = pos;
    g.stp[i] = stp;
    g.len[i] = len;
    g.inb[i] = (pos >= 0.0 && pos < (double)dim) ? 1 : 0;
    g.ipos[i] = (int)pos;
    for (int j = 0; j < W; ++j) {
        double w = 0.0;
        if (j < len) {
            const double r = __dsub_rn((double)(stp + j), pos);
            w = exp(__dmul_rn(__dmul_rn(-1.5, __dmul_rn(r, r)), fsiv));
        }
        g.w[((size_t)a * W + j) * n + d] = w;
    }
}

// z contraction: A[xp][yp][dz] = {sum w s, sum w zp s}
__global__ void k_fv_z(const float* __restrict__ src, AxisGeom g, int n, int W, int sx, int sy, int sz,
                       double2* __restrict__ A) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)sx * sy * n;
    if (idx >= total) return;
    const int dz = (int)(idx % n);
    const size_t line = idx / n;   // xp * sy + yp
    const int st = g.stp[2 * n + dz], len = g.len[2 * n + dz];
    const double* w = g.w + (size_t)2 * W * n + dz;
    const float* s = src + line * sz;
    double a0 = 0.0, a1 = 0.0;
    for (int j = 0; j < len; ++j) {
        const double v = w[(size_t)j * n] * (double)s[st + j];
        a0 += v;
        a1 += v * (double)(st + j);
    }
    A[idx] = make_double2(a0, a1);
}

// y contraction: B[xp][dy][dz] = {sum w A0, sum w yp A0, sum w A1}
__global__ void k_fv_y(const double2* __restrict__ A, AxisGeom g, int n, int W, int sx, int sy,
                       double* __restrict__ B) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)sx * n * n;
    if (idx >= total) return;
    const int dz = (int)(idx % n);
    const int dy = (int)((idx / n) % n);
    const size_t xp = idx / ((size_t)n * n);
    const int st = g.stp[n + dy], len = g.len[n + dy];
    const double* w = g.w + (size_t)W * n + dy;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    for (int j = 0; j < len; ++j) {
        const double2 a = A[(xp * sy + (st + j)) * n + dz];
        const double wj = w[(size_t)j * n];
        const double v = wj * a.x;
        b0 += v;
        b1 += v * (double)(st + j);
        b2 += wj * a.y;
    }
    const size_t plane = (size_t)sx * n * n;
    B[idx] = b0;
    B[plane + idx] = b1;
    B[2 * plane + idx] = b2;
}

// x contraction + the per-voxel tail of doFastVEC (:210-255)
__global__ void k_fv_x(const double* __restrict__ B, const float* __restrict__ src, AxisGeom g, int n,
                       int W, int sx, int sy, int sz, double* __restrict__ data, float* __restrict__ vec) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n3 = (size_t)n * n * n;
    if (idx >= n3) return;
    const int dz = (int)(idx % n);
    const int dy = (int)((idx / n) % n);
    const int dx = (int)(idx / ((size_t)n * n));
    double outd = 0.0;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f;
    const bool inb = g.inb[dx] && g.inb[n + dy] && g.inb[2 * n + dz];
    bool active = false;
    if (inb) {
        const size_t si = ((size_t)g.ipos[dx] * sy + g.ipos[n + dy]) * sz + g.ipos[2 * n + dz];
        active = src[si] != 0.f;
    }
    if (active) {
        const int st = g.stp[dx], len = g.len[dx];
        const double* w = g.w + dx;
        const size_t plane = (size_t)sx * n * n;
        const size_t off = (size_t)dy * n + dz;
        double D = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        for (int j = 0; j < len; ++j) {
            const size_t bi = (size_t)(st + j) * n * n + off;
            const double b0 = B[bi];
            const double wj = w[(size_t)j * n];
            const double v = wj * b0;
            D += v;
            mx += v * (double)(st + j);
            my += wj * B[plane + bi];
            mz += wj * B[2 * plane + bi];
        }
        outd = D;
        if (D != 0.0) {
            const double rd = __ddiv_rn(1.0, D);
            const double t0 = __dsub_rn(__dmul_rn(mx, rd), g.pos[dx]);
            const double t1 = __dsub_rn(__dmul_rn(my, rd), g.pos[n + dy]);
            const double t2 = __dsub_rn(__dmul_rn(mz, rd), g.pos[2 * n + dz]);
            double dv = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(t0, t0), __dmul_rn(t1, t1)), __dmul_rn(t2, t2)));
            if (dv == 0.0) dv = 1.0;
            const double rv = __ddiv_rn(1.0, dv);
            v0 = (float)__dmul_rn(t0, rv);
            v1 = (float)__dmul_rn(t1, rv);
            v2 = (float)__dmul_rn(t2, rv);
        }
    }
    data[idx] = outd;
    vec[3 * idx] = v0;
    vec[3 * idx + 1] = v1;
    vec[3 * idx + 2] = v2;
}

}  // namespace vx

using namespace vx;

extern "C" int vx_fastvec(const float* src, int sx, int sy, int sz, double src_width,
                          const double* so, int n, double dest_width, const double* dorig,
                          double dreso, double* data, float* vec, vx_stream stream) {
    VX_REQUIRE(src && so && dorig && data && vec, VX_ERR_ARG, "NULL argument");
    VX_REQUIRE(sx > 0 && sy > 0 && sz > 0 && n > 0, VX_ERR_ARG, "bad dimensions");
    VX_REQUIRE(src_width > 0 && dreso > 0, VX_ERR_ARG, "bad widths");
    hipStream_t st = (hipStream_t)stream;
    // VESPER_1_prob_new.py:183-187
    const double gstep = src_width;
    double fs = (dreso / gstep) * 0.5;
    fs = fs * fs;
    const double fsiv = 1.0 / fs;
    const double fmaxd = (dreso / gstep) * 2.0;
    const int W = (int)(2.0 * fmaxd) + 3;
    VX_REQUIRE(W > 0 && W < 100000, VX_ERR_ARG, "window too large");
    AxisGeom g;
    char* base = nullptr;
    const size_t nA = (size_t)sx * sy * n, nB = (size_t)sx * n * n;
    const size_t bytes_geom = 3 * (size_t)n * (sizeof(double) + 4 * sizeof(int)) + 3 * (size_t)n * W * sizeof(double);
    const size_t bytes = bytes_geom + 256 + nA * sizeof(double2) + 256 + 3 * nB * sizeof(double);
    VX_HIP(hipMallocAsync((void**)&base, bytes, st));
    char* p = base;
    g.pos = (double*)p; p += 3 * n * sizeof(double);
    g.w = (double*)p; p += 3 * (size_t)n * W * sizeof(double);
    g.stp = (int*)p; p += 3 * n * sizeof(int);
    g.len = (int*)p; p += 3 * n * sizeof(int);
    g.inb = (int*)p; p += 3 * n * sizeof(int);
    g.ipos = (int*)p; p += 3 * n * sizeof(int);
    p = base + ((bytes_geom + 255) / 256) * 256;
    double2* A = (double2*)p; p += ((nA * sizeof(double2) + 255) / 256) * 256;
    double* B = (double*)p;
    hipLaunchKernelGGL(k_fv_axis, dim3((n + 127) / 128, 3), dim3(128), 0, st, g, n, W, sx, sy, sz, dest_width,
                       dorig[0], dorig[1], dorig[2], so[0], so[1], so[2], gstep, fmaxd, fsiv);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fv_z, dim3((unsigned)((nA + 255) / 256)), dim3(256), 0, st, src, g, n, W, sx, sy, sz, A);
    VX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fv_y, dim3((unsigned)((nB + 255) / 256)), dim3(256), 0, st, A, g, n, W, sx, sy, B);
    VX_LAUNCH_CHECK();
    const size_t n3 = (size_t)n * n * n;
    hipLaunchKernelGGL(k_fv_x, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, B, src, g, n, W, sx, sy, sz,
                       data, vec);
    VX_LAUNCH_CHECK();
    VX_HIP(hipFreeAsync(base, st));
    return VX_OK;
}
