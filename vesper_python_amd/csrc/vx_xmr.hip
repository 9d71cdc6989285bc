// x pass with the x-axis DFT factored on the matrix cores (x-pass variant 4).
//
// The correlation's x axis (length N) is split Cooley-Tukey style,
// N = N1 * N2 with N1 <= 32 and N2 in {2, 3, 4, 5, 7, 8}:
//   forward  X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_N^(n2 k1) sum_n1 W_N1^(n1 k1) x[N2 n1 + n2]
//   inverse  y[N2 n1 + n2] = sum_k1 W_N1^-(n1 k1) W_N^-(n2 k1) sum_k2 W_N2^-(n2 k2) P[k1 + N1 k2]
// (W_M = exp(-2 pi i / M)).  The length-N1 DFTs run as complex matrix
// products on v_mfma_f32_16x16x32_f16 with split half-precision operands
// (v = hi + lo, hi = f16(v), lo = f16(v - hi); hi*hi + hi*lo + lo*hi summed in
// fp32, so each product carries ~22 significant bits); each tile's operands
// are scaled by a power of two so that they sit inside f16's range.  The
// twiddles, the length-N2 DFTs, the product with the conjugated target
// spectra and the channel sum stay in fp32 registers: the MFMA result layout
// gives every lane all N2 values of its (k1, column), so the radix-N2 stage
// needs no LDS round trip.  LDS holds only the MFMA data operands (a double
// buffer, one barrier pair per channel).
//
// Workgroup = 4 waves on one tile of 16 flattened (ky, kz) columns of one
// rotation.  Wave v owns complex DFT rows 8v .. 8v + 7 (MFMA rows interleave
// Re / Im), so lane (g = lane / 16, t = lane % 16) of wave v holds rows
// k1 = 8v + 2g + s (s = 0, 1) of column t in every N-tile n2.
// 16x16x32 operand layout: lane l holds A[l % 16][k] and B[k][l % 16] for
// k = 8 (l / 16) + u, u < 8 (any fixed k assignment shared by both operands
// gives the product), and D[4 (l / 16) + v][l % 16], v < 4.  Here k = 8g + u
// stands for complex index 4g + u / 2 of the K-step, Re for even u, Im for
// odd u, so a lane's B operand is four consecutive complex values.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <algorithm>
#include <string.h>
#include <vector>

#include "vx_common.hpp"
#include "vx_fft.hpp"
#include "vx_internal.hpp"

namespace vx {

typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMrT = 16;          // columns per tile
constexpr int kMrMaxN = 256;      // N1 <= 32, N2 <= 8
constexpr int kMrHeadroom = 14;   // scaled operands stay below 2^14 (f16 max 65504)

// (re, im) -> packed f16 (hi) and the packed f16 remainders (lo)
__device__ __forceinline__ void split_pair(float re, float im, uint32_t& hi, uint32_t& lo) {
    const _Float16 hr = (_Float16)re, hm = (_Float16)im;
    const f16x2_t h = {hr, hm};
    const f16x2_t l = {(_Float16)(re - (float)hr), (_Float16)(im - (float)hm)};
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}

__device__ __forceinline__ void split_val(float v, _Float16& hi, _Float16& lo) {
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// Power-of-two scale bringing the tile's largest |component| below 2^14,
// and its inverse (exact).
__device__ __forceinline__ void tile_scale(float mx, float& sc, float& isc) {
    if (!(mx > 0.f) || !(mx < 3.0e38f)) { sc = 1.f; isc = 1.f; return; }
    int e;
    (void)frexpf(mx, &e);   // mx < 2^e
    sc = ldexpf(1.f, kMrHeadroom - e);
    isc = ldexpf(1.f, e - kMrHeadroom);
}

__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

template <int N2, int KSF, int KSI, int CH>
__global__ void __launch_bounds__(256)
k_xcorr_mr(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
           const float2* __restrict__ wtab, const uint4* __restrict__ atab, int n, int N1, int nhz, int lo, int w,
           int nlo, int C, int nb, int ntiles) {
    // LDS: W_N^j (n entries) | 2 x 4 channel maxima | operand buffers
    extern __shared__ float2 mr_lds[];
    float2* sw = mr_lds;
    float* red = (float*)(mr_lds + kMrMaxN);                  // [2][4]
    uint32_t* ops = (uint32_t*)(mr_lds + kMrMaxN + 4);        // forward: [2][2][KSF][N2][64][4]; inverse aliases
    constexpr int kFwdWords = KSF * N2 * 64 * 4;              // one hi (or lo) plane of one buffer
    constexpr int kInvWords = KSI * N2 * 64 * 4;
    constexpr int kPer = KSF * N2;                            // input values per thread and channel

    int b, tile;
    if (!xcd_tile(blockIdx.x, nb, ntiles, &tile, &b)) return;
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, g = l >> 4, t = l & 15;
    const int plane = n * nhz;
    const int j0 = tile * kMrT;
    const bool colok = j0 + t < plane;

    // A operands (xcorr_mr_tables): fragment-ordered, read where used (they
    // stay in L1/L2; holding them would cost 8 (KSF + KSI) registers)
    const uint4* afw = atab + (size_t)wv * KSF * 128 + l;
    const uint4* aiw = atab + (size_t)4 * KSF * 128 + (size_t)wv * KSI * 128 + l;

    // this thread's input elements: n1 = nlo + m, m = 4 wv + g + 16 kk (kk <
    // KSF), every n2 (row x = N2 n1 + n2), column t.  A wave's 64 lanes then
    // write one operand block, lane t + 16 wv, pair g: 64 distinct banks.
    const int mq = 4 * wv + g;
    // Loads without branches, through buffer resources: a channel's input
    // resource spans rows [lo, lo + w) of the tile's columns, so rows outside
    // the support box read as zero (a row below lo gives a negative offset,
    // which wraps past the range); lanes past the plane's last column read
    // column j0 and are zeroed by `cm`.  Every offset is a 32-bit VGPR.
    const int tcol = colok ? t : 0;
    const float cm = colok ? 1.f : 0.f;
    auto load_channel = [&](int c, float2 (&xv)[kPer]) {
        const auto rs = buf_rsrc(spec + (((size_t)b * C + c) * n + lo) * (size_t)plane + j0,
                                (uint32_t)w * plane * sizeof(float2));
        const int r0 = N2 * (nlo + mq) - lo;   // may be negative
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int r = r0 + N2 * 16 * (i / N2) + i % N2;
            const float2 v = buf_load_f2(rs, (uint32_t)((r * plane + tcol) * (int)sizeof(float2)));
            xv[i] = make_float2(v.x * cm, v.y * cm);
        }
    };

    // rows k1 + N1 k2 of a channel's conjugated target at this lane's column
    // (rows past the channel read as zero; rows of k1 >= N1 are never used)
    auto load_target = [&](int c, float2 (&tv)[2][N2]) {
        const auto rt = buf_rsrc(tgt + (size_t)c * n * plane + j0, (uint32_t)(n * plane - j0) * sizeof(float2));
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int k1 = 8 * wv + 2 * g + s;
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2)
                tv[s][k2] = buf_load_f2(rt, (uint32_t)(((k1 + N1 * k2) * plane + tcol) * (int)sizeof(float2)));
        }
    };

    // Every channel's input is loaded up front and the tile takes one scale
    // for all of them, so the channel sum accumulates in the scaled domain
    // and the scale comes off once, with the inverse's.
    float2 xv[CH][kPer];
#pragma unroll
    for (int c = 0; c < CH; ++c) load_channel(c, xv[c]);
    // the twiddle table fills LDS while the inputs are in flight (the
    // barrier below publishes both it and the maxima)
    for (int j = tid; j < n; j += 256) sw[j] = wtab[j];
    float sc, isc;
    {
        float mx = 0.f;
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int i = 0; i < kPer; ++i) mx = fmaxf(mx, fmaxf(fabsf(xv[c][i].x), fabsf(xv[c][i].y)));
        mx = wave_max(mx);
        if (l == 0) red[wv] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        tile_scale(mx, sc, isc);
    }
    float2 P[2][N2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int k = 0; k < N2; ++k) P[s][k] = make_float2(0.f, 0.f);

#pragma unroll
    for (int c = 0; c < CH; ++c) {
        // operand buffer c % 2 was last read by channel c - 2's products,
        // which every wave finished before channel c - 1's barrier
        uint32_t* bh = ops + (c & 1) * 2 * kFwdWords;
        uint32_t* bl = bh + kFwdWords;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int ks = i / N2, n2 = i % N2;
            const int wi = (((ks * N2 + n2) * 64 + t + 16 * wv) << 2) + g;
            uint32_t h, lw;
            split_pair(xv[c][i].x * sc, xv[c][i].y * sc, h, lw);
            bh[wi] = h;
            bl[wi] = lw;
        }
        // the target's loads are in flight while the waves meet at the barrier
        float2 tv[2][N2];
        load_target(c, tv);
        __syncthreads();
        f32x4_t acc[N2];
#pragma unroll
        for (int k = 0; k < N2; ++k) acc[k] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KSF; ++ks) {
            const f16x8_t afh = __builtin_bit_cast(f16x8_t, afw[ks * 128]);
            const f16x8_t afl = __builtin_bit_cast(f16x8_t, afw[ks * 128 + 64]);
#pragma unroll
            for (int k = 0; k < N2; ++k) {
                const int wi = ((ks * N2 + k) * 64 + l) << 2;
                const f16x8_t xh = __builtin_bit_cast(f16x8_t, *(const uint4*)(bh + wi));
                const f16x8_t xl = __builtin_bit_cast(f16x8_t, *(const uint4*)(bl + wi));
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(afh, xh, acc[k], 0, 0, 0);
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(afh, xl, acc[k], 0, 0, 0);
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(afl, xh, acc[k], 0, 0, 0);
            }
        }
        // twiddle, length-N2 DFT, times the target, channel sum (all scaled by sc)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int k1 = 8 * wv + 2 * g + s;
            if (k1 < N1) {
                float2 u[N2];
#pragma unroll
                for (int k = 0; k < N2; ++k) {
                    u[k] = make_float2(acc[k][2 * s], acc[k][2 * s + 1]);
                    if (k > 0) u[k] = cmul(u[k], sw[k * k1]);
                }
                dft<N2>(u);
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) P[s][k2] = cadd(P[s][k2], cmul(u[k2], tv[s][k2]));
            }
        }
    }

    // inverse: length-N2 DFTs and twiddles in registers, then the N1 DFTs
    float2 R[2][N2];
    float mx = 0.f;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int k1 = 8 * wv + 2 * g + s;
#pragma unroll
        for (int k = 0; k < N2; ++k) R[s][k] = cconj(P[s][k]);
        dft<N2>(R[s]);
#pragma unroll
        for (int k = 0; k < N2; ++k) {
            R[s][k] = cconj(R[s][k]);
            if (k > 0 && k1 < N1) R[s][k] = cmulc(R[s][k], sw[k * k1]);
            mx = fmaxf(mx, fmaxf(fabsf(R[s][k].x), fabsf(R[s][k].y)));
        }
    }
    mx = wave_max(mx);
    if (l == 0) red[4 + wv] = mx;
    __syncthreads();   // also: every wave is done with the forward operand buffers
    mx = fmaxf(fmaxf(red[4], red[5]), fmaxf(red[6], red[7]));
    float sc2, isc2;
    tile_scale(mx, sc2, isc2);
    const float osc = isc2 * isc;   // both powers of two: exact
    uint32_t* qh = ops;
    uint32_t* ql = ops + kInvWords;
    if (wv < 2 * KSI) {
        // k1 = 8 wv + 2 g + s: K-step wv / 2, lane t + 16 (2 (wv % 2) + g / 2), pair 2 (g % 2) + s
        const int ks = wv >> 1, gp = 2 * (wv & 1) + (g >> 1), jp = 2 * (g & 1);
#pragma unroll
        for (int k = 0; k < N2; ++k) {
            uint32_t h0, l0, h1, l1;
            split_pair(R[0][k].x * sc2, R[0][k].y * sc2, h0, l0);
            split_pair(R[1][k].x * sc2, R[1][k].y * sc2, h1, l1);
            const int wi = (((ks * N2 + k) * 64 + t + 16 * gp) << 2) + jp;
            *(uint2*)(qh + wi) = make_uint2(h0, h1);
            *(uint2*)(ql + wi) = make_uint2(l0, l1);
        }
    }
    __syncthreads();
    f32x4_t acc[N2];
#pragma unroll
    for (int k = 0; k < N2; ++k) acc[k] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KSI; ++ks) {
        const f16x8_t aih = __builtin_bit_cast(f16x8_t, aiw[ks * 128]);
        const f16x8_t ail = __builtin_bit_cast(f16x8_t, aiw[ks * 128 + 64]);
#pragma unroll
        for (int k = 0; k < N2; ++k) {
            const int wi = ((ks * N2 + k) * 64 + l) << 2;
            const f16x8_t xh = __builtin_bit_cast(f16x8_t, *(const uint4*)(qh + wi));
            const f16x8_t xl = __builtin_bit_cast(f16x8_t, *(const uint4*)(ql + wi));
            acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, xh, acc[k], 0, 0, 0);
            acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, xl, acc[k], 0, 0, 0);
            acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, xh, acc[k], 0, 0, 0);
        }
    }
    float2* dst = out + (size_t)b * n * (size_t)plane + j0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int n1 = 8 * wv + 2 * g + s;
        if (colok && n1 < N1) {
#pragma unroll
            for (int k = 0; k < N2; ++k)
                dst[(uint32_t)((N2 * n1 + k) * plane + t)] = make_float2(acc[k][2 * s] * osc, acc[k][2 * s + 1] * osc);
        }
    }
}

// N = N1 N2 with N2 the smallest supported radix that leaves N1 <= 32.
bool xcorr_mr_factor(int n, int* n1, int* n2) {
    static const int radices[] = {2, 3, 4, 5, 7, 8};
    for (int r : radices)
        if (n % r == 0 && n / r <= 32 && n / r >= 2) {
            *n1 = n / r;
            *n2 = r;
            return true;
        }
    return false;
}

static void mr_shape(int n, int lo, int w, int* N1, int* N2, int* nlo, int* ksf, int* ksi) {
    xcorr_mr_factor(n, N1, N2);
    *nlo = lo / *N2;
    const int nwin = (lo + w - 1) / *N2 - *nlo + 1;   // <= N1 <= 32
    *ksf = nwin <= 16 ? 1 : 2;
    *ksi = *N1 <= 16 ? 1 : 2;
}

static uint16_t f16_bits(float v) {
    const _Float16 h = (_Float16)v;
    uint16_t u;
    memcpy(&u, &h, 2);
    return u;
}

// The A operands of k_xcorr_mr in fragment order, as 16-byte lane records:
// forward [wave 4][K-step KSF][hi, lo][lane 64], then inverse [4][KSI][2][64].
// Forward rows (k1 = 8 wave + t / 2, Re / Im by t % 2), columns (n1 = nlo +
// 16 ks + 4 g + u / 2, Re / Im by u % 2), e^{-i theta}; inverse rows (n1 =
// 8 wave + t / 2), columns (k1 = 16 ks + 4 g + u / 2), e^{+i theta};
// theta = 2 pi (k1 n1 mod N1) / N1, zero outside N1.
int xcorr_mr_tables(int n, int lo, int w, std::vector<uint32_t>& out) {
    int N1, N2, nlo, ksf, ksi;
    VX_REQUIRE(xcorr_mr_factor(n, &N1, &N2), VX_ERR_UNSUPPORTED, "grid edge %d has no MFMA radix split", n);
    mr_shape(n, lo, w, &N1, &N2, &nlo, &ksf, &ksi);
    out.assign((size_t)4 * (ksf + ksi) * 2 * 64 * 4, 0u);
    auto put = [&](size_t rec, int u, float v) {
        const _Float16 h = (_Float16)v;
        const float lo_ = v - (float)h;
        uint16_t* hh = (uint16_t*)&out[rec * 4];
        uint16_t* ll = (uint16_t*)&out[(rec + 64) * 4];
        hh[u] = f16_bits(v);
        ll[u] = f16_bits(lo_);
    };
    for (int fwd = 1; fwd >= 0; --fwd) {
        const int KS = fwd ? ksf : ksi;
        const size_t base = fwd ? 0 : (size_t)4 * ksf * 128;
        for (int wv = 0; wv < 4; ++wv)
            for (int ks = 0; ks < KS; ++ks)
                for (int l = 0; l < 64; ++l) {
                    const int t = l & 15, g = l >> 4;
                    const int rk = 8 * wv + (t >> 1), ro = t & 1;
                    for (int u = 0; u < 8; ++u) {
                        const int col = fwd ? nlo + 16 * ks + 4 * g + (u >> 1) : 16 * ks + 4 * g + (u >> 1);
                        const int ri = u & 1;
                        float v = 0.f;
                        if (rk < N1 && col < N1) {
                            const double th = 2.0 * M_PI * (double)((rk * col) % N1) / N1;
                            const float c = (float)cos(th), sn = (float)sin(th);
                            if (fwd) v = ro == 0 ? (ri == 0 ? c : sn) : (ri == 0 ? -sn : c);
                            else     v = ro == 0 ? (ri == 0 ? c : -sn) : (ri == 0 ? sn : c);
                        }
                        put(base + ((size_t)wv * KS + ks) * 128 + l, u, v);
                    }
                }
    }
    return VX_OK;
}

size_t xcorr_mr_lds(int n2, int ksf, int ksi) {
    const size_t ops = (size_t)std::max(4 * ksf, 2 * ksi) * n2 * 64 * 4 * sizeof(uint32_t);
    return (kMrMaxN + 4) * sizeof(float2) + ops;
}

template <int N2, int KSF, int KSI, int CH>
static int launch_mr_t(dim3 grid, hipStream_t st, const float2* spec, const float2* tgt, float2* out,
                       const float2* wtab, const uint4* atab, int n, int N1, int nhz, int lo, int w, int nlo, int C,
                       int nb, int ntiles) {
    const size_t lds = xcorr_mr_lds(N2, KSF, KSI);
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_xcorr_mr<N2, KSF, KSI, CH>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_xcorr_mr<N2, KSF, KSI, CH>), grid, dim3(256), lds, st, spec, tgt, out, wtab, atab, n, N1, nhz,
                       lo, w, nlo, C, nb, ntiles);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

template <int N2, int CH>
static int launch_mr_c(int ksf, int ksi, dim3 grid, hipStream_t st, const float2* spec, const float2* tgt,
                       float2* out, const float2* wtab, const uint4* atab, int n, int N1, int nhz, int lo, int w,
                       int nlo, int C, int nb, int ntiles) {
    if (ksf == 1 && ksi == 1) return launch_mr_t<N2, 1, 1, CH>(grid, st, spec, tgt, out, wtab, atab, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
    if (ksf == 1 && ksi == 2) return launch_mr_t<N2, 1, 2, CH>(grid, st, spec, tgt, out, wtab, atab, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
    if (ksf == 2 && ksi == 1) return launch_mr_t<N2, 2, 1, CH>(grid, st, spec, tgt, out, wtab, atab, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
    return launch_mr_t<N2, 2, 2, CH>(grid, st, spec, tgt, out, wtab, atab, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
}

template <int N2>
static int launch_mr_r(int ksf, int ksi, dim3 grid, hipStream_t st, const float2* spec, const float2* tgt,
                       float2* out, const float2* wtab, const uint4* atab, int n, int N1, int nhz, int lo, int w,
                       int nlo, int C, int nb, int ntiles) {
    if (C == 3) return launch_mr_c<N2, 3>(ksf, ksi, grid, st, spec, tgt, out, wtab, atab, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
    return launch_mr_c<N2, 1>(ksf, ksi, grid, st, spec, tgt, out, wtab, atab, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
}

int launch_xcorr_mr(const float2* spec, const float2* tgt, float2* out, const float2* wtab, const void* atab, int n,
                    int nb, int C, int lo, int w, hipStream_t st) {
    int N1, N2, nlo, ksf, ksi;
    VX_REQUIRE(xcorr_mr_factor(n, &N1, &N2), VX_ERR_UNSUPPORTED,
               "grid edge %d has no factor N1 * N2 with N1 <= 32 for the MFMA x pass", n);
    VX_REQUIRE(lo >= 0 && w >= 1 && lo + w <= n, VX_ERR_STATE, "bad support box");
    VX_REQUIRE(C == 1 || C == 3, VX_ERR_UNSUPPORTED, "the MFMA radix-split x pass takes 1 or 3 channels");
    mr_shape(n, lo, w, &N1, &N2, &nlo, &ksf, &ksi);
    const int nhz = n / 2 + 1;
    const int ntiles = (n * nhz + kMrT - 1) / kMrT;
    const dim3 grid((unsigned)(((ntiles + 7) / 8) * 8 * nb));
    const uint4* a = (const uint4*)atab;
    switch (N2) {
        case 2: return launch_mr_r<2>(ksf, ksi, grid, st, spec, tgt, out, wtab, a, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
        case 3: return launch_mr_r<3>(ksf, ksi, grid, st, spec, tgt, out, wtab, a, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
        case 4: return launch_mr_r<4>(ksf, ksi, grid, st, spec, tgt, out, wtab, a, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
        case 5: return launch_mr_r<5>(ksf, ksi, grid, st, spec, tgt, out, wtab, a, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
        case 7: return launch_mr_r<7>(ksf, ksi, grid, st, spec, tgt, out, wtab, a, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
        default: return launch_mr_r<8>(ksf, ksi, grid, st, spec, tgt, out, wtab, a, n, N1, nhz, lo, w, nlo, C, nb, ntiles);
    }
}

}  // namespace vx
