// LDS-resident mixed-radix FFT engine for CDNA4 (gfx950).
//
// A workgroup holds `nlines` complex lines of length n in LDS and transforms
// them together with self-sorting (Stockham) stages, radices 2/3/4/5/7/8/11/13
// -- every length 2^a 3^b 5^c 7^d 11^e 13^f that pyFFTW's next_fast_len can
// return (the grid edge of the reference, VESPER_1_prob_new.py:94-97).
//
// Stage with radix R after p = product of earlier radices (T = n / R):
//   butterfly i in [0, T):  k = i mod p
//     x_j = in[i + j T] * w^(j k),  w = exp(-2 pi i / (p R))
//     y   = DFT_R(x)
//     out[(i - k) R + k + j p] = y_j
// Every thread first reads and transforms all of its butterflies into
// registers, the block synchronises, then all threads write: stages are
// in place in one LDS buffer.  Forward transforms only; the inverse
// (unnormalised, FFTW_BACKWARD) is conj(FFT(conj(x))), applied by callers at
// their loads/stores.
//
// The transform arithmetic below may contract a*b + c into one fma (it only
// has to meet the search's 1e-4 score tolerance, and it stays deterministic);
// the file restores the translation unit's no-contraction default at its end,
// which the parity-critical gather and mixing code relies on.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(fast)

namespace vx {

constexpr int kMaxStages = 16;

// Per-stage parameters live in device memory (read with uniform scalar
// loads); only the pointers and sizes travel as kernel arguments.
struct FftStage {
    int R;          // radix
    int P;          // product of the earlier radices
    float Pinv;     // 1 / P, for the in-kernel k = i mod P
    int tw;         // float2 offset of the stage's twiddles, laid out [j - 1][k], j < R, k < P
};

struct FftDesc {
    int n;
    int nst;
    int ntw;                     // twiddle table length (float2)
    unsigned radix_mask;         // bit r set when a radix-r stage occurs
    unsigned long long rcode;    // stage radices in order, 4 bits each (stage 0 lowest)
    const FftStage* st;          // device [nst]
    const float2* twid;          // twiddle table: device, or LDS once staged
};

// Copy the twiddle table into LDS (all threads of the block take part; the
// caller's next __syncthreads publishes it) and return a descriptor reading it.
__device__ __forceinline__ FftDesc stage_twiddles(const FftDesc& d, float2* lds_tw) {
    for (int i = threadIdx.x; i < d.ntw; i += blockDim.x) lds_tw[i] = d.twid[i];
    FftDesc r = d;
    r.twid = lds_tw;
    return r;
}

// Quotient of small non-negative integers (a < 2^22, d < 2^16) through a
// float reciprocal, corrected to exact: avoids the ~40-instruction integer
// division in butterfly index math.
// Index products below fit 24 bits (offsets < 2^22): the full-rate 24-bit
// multiply replaces the quarter-rate 32-bit one.
__device__ __forceinline__ int imul(int a, int b) { return __mul24(a, b); }

__device__ __forceinline__ int fdiv(int a, int d, float dinv) {
    int q = (int)((float)a * dinv);
    const int r = a - imul(q, d);
    q += (r >= d) ? 1 : ((r < 0) ? -1 : 0);
    return q;
}

// a / d for small non-negative a: an unsigned constant-divisor divide (a
// multiply-shift) when the divisor is a compile-time constant (S), else the
// float-reciprocal form above.
template <bool S>
__device__ __forceinline__ int qdiv(int a, int d, float dinv) {
    if constexpr (S) return (int)((unsigned)a / (unsigned)d);
    else return fdiv(a, d, dinv);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// -i * a
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

template <int R> struct OddRadix;
template <> struct OddRadix<3> {
  static constexpr float C[3] = {1.000000000e+00f, -5.000000000e-01f, -5.000000000e-01f};
  static constexpr float S[3] = {0.000000000e+00f, 8.660254038e-01f, -8.660254038e-01f};
};
template <> struct OddRadix<5> {
  static constexpr float C[5] = {1.000000000e+00f, 3.090169944e-01f, -8.090169944e-01f, -8.090169944e-01f, 3.090169944e-01f};
  static constexpr float S[5] = {0.000000000e+00f, 9.510565163e-01f, 5.877852523e-01f, -5.877852523e-01f, -9.510565163e-01f};
};
template <> struct OddRadix<7> {
  static constexpr float C[7] = {1.000000000e+00f, 6.234898019e-01f, -2.225209340e-01f, -9.009688679e-01f, -9.009688679e-01f, -2.225209340e-01f, 6.234898019e-01f};
  static constexpr float S[7] = {0.000000000e+00f, 7.818314825e-01f, 9.749279122e-01f, 4.338837391e-01f, -4.338837391e-01f, -9.749279122e-01f, -7.818314825e-01f};
};
template <> struct OddRadix<11> {
  static constexpr float C[11] = {1.000000000e+00f, 8.412535328e-01f, 4.154150130e-01f, -1.423148383e-01f, -6.548607339e-01f, -9.594929736e-01f, -9.594929736e-01f, -6.548607339e-01f, -1.423148383e-01f, 4.154150130e-01f, 8.412535328e-01f};
  static constexpr float S[11] = {0.000000000e+00f, 5.406408175e-01f, 9.096319954e-01f, 9.898214419e-01f, 7.557495744e-01f, 2.817325568e-01f, -2.817325568e-01f, -7.557495744e-01f, -9.898214419e-01f, -9.096319954e-01f, -5.406408175e-01f};
};
template <> struct OddRadix<13> {
  static constexpr float C[13] = {1.000000000e+00f, 8.854560257e-01f, 5.680647467e-01f, 1.205366803e-01f, -3.546048870e-01f, -7.485107482e-01f, -9.709418174e-01f, -9.709418174e-01f, -7.485107482e-01f, -3.546048870e-01f, 1.205366803e-01f, 5.680647467e-01f, 8.854560257e-01f};
  static constexpr float S[13] = {0.000000000e+00f, 4.647231720e-01f, 8.229838659e-01f, 9.927088741e-01f, 9.350162427e-01f, 6.631226582e-01f, 2.393156643e-01f, -2.393156643e-01f, -6.631226582e-01f, -9.350162427e-01f, -9.927088741e-01f, -8.229838659e-01f, -4.647231720e-01f};
};

// Forward DFT of odd length R (exp(-2 pi i jk / R)), folded over j <-> R-j.
template <int R>
__device__ __forceinline__ void dft_odd(float2 (&x)[R]) {
    constexpr int H = (R - 1) / 2;
    float2 s[H], d[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        s[j - 1] = cadd(x[j], x[R - j]);
        d[j - 1] = csub(x[j], x[R - j]);
    }
    float2 y[R];
    float2 y0 = x[0];
#pragma unroll
    for (int j = 0; j < H; ++j) y0 = cadd(y0, s[j]);
    y[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        float2 A = x[0];
        float2 B = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const float c = OddRadix<R>::C[(j * k) % R];
            const float sn = OddRadix<R>::S[(j * k) % R];
            A.x += s[j - 1].x * c;  A.y += s[j - 1].y * c;
            B.x += d[j - 1].x * sn; B.y += d[j - 1].y * sn;
        }
        y[k] = make_float2(A.x + B.y, A.y - B.x);
        y[R - k] = make_float2(A.x - B.y, A.y + B.x);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = y[j];
}

__device__ __forceinline__ void dft2(float2 (&x)[2]) {
    float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

__device__ __forceinline__ void dft4(float2 (&x)[4]) {
    float2 a0 = cadd(x[0], x[2]), b0 = csub(x[0], x[2]);
    float2 a1 = cadd(x[1], x[3]), b1 = cmul_mi(csub(x[1], x[3]));
    x[0] = cadd(a0, a1);
    x[2] = csub(a0, a1);
    x[1] = cadd(b0, b1);
    x[3] = csub(b0, b1);
}

__device__ __forceinline__ void dft8(float2 (&x)[8]) {
    constexpr float r = 7.071067812e-01f;
    float2 a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        a[j] = cadd(x[j], x[j + 4]);
        b[j] = csub(x[j], x[j + 4]);
    }
    // b_j *= exp(-2 pi i j / 8)
    b[1] = make_float2((b[1].x + b[1].y) * r, (b[1].y - b[1].x) * r);
    b[2] = cmul_mi(b[2]);
    b[3] = make_float2((b[3].y - b[3].x) * r, -(b[3].x + b[3].y) * r);
    float2 e[4] = {a[0], a[1], a[2], a[3]};
    float2 o[4] = {b[0], b[1], b[2], b[3]};
    dft4(e);
    dft4(o);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x[2 * k] = e[k];
        x[2 * k + 1] = o[k];
    }
}

template <int R>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
    if constexpr (R == 2) dft2(x);
    else if constexpr (R == 4) dft4(x);
    else if constexpr (R == 8) dft8(x);
    else dft_odd<R>(x);
}

// One Stockham stage over all lines held in LDS.
//   buf:  LDS base;  element e of line l lives at buf[l * ls + e * es]
//   line_fast: map consecutive threads to consecutive lines (use when the
//              line stride is 1) instead of consecutive butterflies.
template <int R, int MAXV>
__device__ __forceinline__ void fft_stage(float2* buf, int n, int nlines, int es, int ls,
                                          bool line_fast, int p, float pinv, const float2* __restrict__ tw) {
    constexpr int MAXB = (MAXV + R - 1) / R;
    const int T = n / R;
    const int nb = T * nlines;
    const int nthr = blockDim.x;
    const float linv = 1.0f / (float)nlines, tinv = 1.0f / (float)T;
    const int lg = (nlines & (nlines - 1)) == 0 ? __ffs(nlines) - 1 : -1;
    float2 v[MAXB][R];
    int obase[MAXB];
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        const int b = threadIdx.x + m * nthr;
        obase[m] = -1;
        if (b < nb) {
            int line, i;
            if (line_fast && lg >= 0) { i = b >> lg; line = b & (nlines - 1); }
            else if (line_fast) { i = fdiv(b, nlines, linv); line = b - imul(i, nlines); }
            else           { line = fdiv(b, T, tinv);   i = b - imul(line, T); }
            float2* L = buf + imul(line, ls) + imul(i, es);
            const int tes = imul(T, es);
#pragma unroll
            for (int j = 0; j < R; ++j) v[m][j] = L[j * tes];
            const int k = (p == 1) ? 0 : (i - imul(fdiv(i, p, pinv), p));
            if (k != 0) {
                const float2* w = tw + k;
#pragma unroll
                for (int j = 1; j < R; ++j) v[m][j] = cmul(v[m][j], w[(j - 1) * p]);
            }
            dft<R>(v[m]);
            obase[m] = imul(line, ls) + imul(imul(i - k, R) + k, es);
        }
    }
    __syncthreads();
    const int pes = imul(p, es);
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        if (obase[m] >= 0) {
#pragma unroll
            for (int j = 0; j < R; ++j) buf[obase[m] + j * pes] = v[m][j];
        }
    }
    __syncthreads();
}

// Forward FFT of `nlines` lines in LDS (caller has synchronised after filling).
template <int MAXV, unsigned RM = 0xffffffffu>
__device__ void fft_lds(float2* buf, const FftDesc& d, int nlines, int es, int ls, bool line_fast) {
    for (int s = 0; s < d.nst; ++s) {
        const FftStage sg = d.st[s];
        const float2* tw = d.twid + sg.tw;
        const int p = sg.P;
        const float pi = sg.Pinv;
        switch (sg.R) {
#define VX_IP_CASE(r) case r: if constexpr ((RM & (1u << r)) != 0) fft_stage<r, MAXV>(buf, d.n, nlines, es, ls, line_fast, p, pi, tw); break;
            VX_IP_CASE(2) VX_IP_CASE(3) VX_IP_CASE(4) VX_IP_CASE(5)
            VX_IP_CASE(7) VX_IP_CASE(8) VX_IP_CASE(11) VX_IP_CASE(13)
#undef VX_IP_CASE
            default: break;
        }
    }
}

// Ping-pong variant: each stage reads `in` and writes `out` (two LDS
// buffers), so a thread keeps only one butterfly in registers at a time.
// Ends synchronised; returns the buffer that holds the result.
template <int R>
__device__ __forceinline__ void butterfly_pp(const float2* __restrict__ in, float2* __restrict__ out, int b, int T,
                                             int nlines, float linv, float tinv, int es, int ls, bool line_fast,
                                             int p, float pinv, const float2* __restrict__ tw, int lg) {
    int line, i;
    if (line_fast && lg >= 0) { i = b >> lg; line = b & (nlines - 1); }   // power-of-two line count
    else if (line_fast) { i = fdiv(b, nlines, linv); line = b - imul(i, nlines); }
    else           { line = fdiv(b, T, tinv);   i = b - imul(line, T); }
    const float2* L = in + imul(line, ls) + imul(i, es);
    const int tes = imul(T, es);
    float2 v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = L[j * tes];
    const int k = (p == 1) ? 0 : (i - imul(fdiv(i, p, pinv), p));
    if (k != 0) {
        const float2* w = tw + k;
#pragma unroll
        for (int j = 1; j < R; ++j) v[j] = cmul(v[j], w[(j - 1) * p]);
    }
    dft<R>(v);
    float2* O = out + imul(line, ls) + imul(imul(i - k, R) + k, es);
    const int pes = imul(p, es);
#pragma unroll
    for (int j = 0; j < R; ++j) O[j * pes] = v[j];
}

template <int R>
__device__ __forceinline__ void fft_stage_pp(const float2* __restrict__ in, float2* __restrict__ out, int n,
                                             int nlines, int es, int ls, bool line_fast, int p, float pinv,
                                             const float2* __restrict__ tw) {
    const int T = n / R;
    const int nb = T * nlines;
    const int nthr = blockDim.x;
    const float linv = 1.0f / (float)nlines, tinv = 1.0f / (float)T;
    const int lg = (nlines & (nlines - 1)) == 0 ? __ffs(nlines) - 1 : -1;
    for (int b = threadIdx.x; b < nb; b += nthr)
        butterfly_pp<R>(in, out, b, T, nlines, linv, tinv, es, ls, line_fast, p, pinv, tw, lg);
    __syncthreads();
}

// Radix sets: bit r of a mask enables the radix-r stage code in a kernel
// (kernels instantiated for the radices of their plan's length carry less
// code and scalar state); kAllRadices compiles all of them.
constexpr unsigned radix_bit(int r) { return 1u << r; }
constexpr unsigned kAllRadices = radix_bit(2) | radix_bit(3) | radix_bit(4) | radix_bit(5) | radix_bit(7) |
                                 radix_bit(8) | radix_bit(11) | radix_bit(13);

template <unsigned RM = kAllRadices>
__device__ __forceinline__ float2* fft_lds_pp(float2* a, float2* b, const FftDesc& d, int nlines, int es, int ls,
                                              bool line_fast) {
    for (int s = 0; s < d.nst; ++s) {
        const FftStage sg = d.st[s];
        const float2* tw = d.twid + sg.tw;
        const int p = sg.P;
        const float pi = sg.Pinv;
        switch (sg.R) {
#define VX_PP_CASE(r) case r: if constexpr ((RM & radix_bit(r)) != 0) fft_stage_pp<r>(a, b, d.n, nlines, es, ls, line_fast, p, pi, tw); break;
            VX_PP_CASE(2) VX_PP_CASE(3) VX_PP_CASE(4) VX_PP_CASE(5)
            VX_PP_CASE(7) VX_PP_CASE(8) VX_PP_CASE(11) VX_PP_CASE(13)
#undef VX_PP_CASE
            default: break;
        }
        float2* t = a; a = b; b = t;
    }
    return a;
}

// ---------------------------------------------------------------------------
// Compile-time column-tile FFTs.
//
// A tile holds NL lines of length n as [element][line] (element stride NL,
// line stride 1) and consecutive threads take consecutive lines, the layout
// of the x pass.  Butterfly b = line + i NL then reads in[b + j T NL] and
// writes out[b + (R - 1) NL (i - k) + j P NL].  When the tile width and the
// plan's radix order are known at compile time (StaticTile<NL, R0, R1, ...>,
// radices in the order fft_desc_create chooses), every stage's R, P, span T,
// twiddle offset and butterfly count are constants: the per-thread butterfly
// loop unrolls, i = b / NL and k = i mod P reduce to shifts and
// multiply-shifts, the LDS offsets become instruction immediates, and the
// twiddle multiply runs unconditionally for P > 1 (the k = 0 twiddles are
// exactly 1, so the values match the runtime form).
// ---------------------------------------------------------------------------

template <int N, int NL, int NT, int R, int P, int TW>
__device__ __forceinline__ void tile_stage_static(const float2* __restrict__ in, float2* __restrict__ out,
                                                  const float2* __restrict__ tw) {
    constexpr int T = N / R;
    constexpr int NB = T * NL;
#pragma unroll 1
    for (int b0 = 0; b0 < NB; b0 += NT) {
        const int b = b0 + (int)threadIdx.x;
        if (NB % NT == 0 || b < NB) {
            const int i = b / NL;
            float2 v[R];
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = in[b + j * T * NL];
            int k = 0;
            if constexpr (P > 1) {
                k = (int)((unsigned)i % (unsigned)P);
                const float2* w = tw + TW + k;
#pragma unroll
                for (int j = 1; j < R; ++j) v[j] = cmul(v[j], w[(j - 1) * P]);
            }
            dft<R>(v);
            float2* O = out + b + (R - 1) * NL * (i - k);
#pragma unroll
            for (int j = 0; j < R; ++j) O[j * P * NL] = v[j];
        }
    }
    __syncthreads();
}

template <int N, int NL, int NT, int P, int TW, int R, int... Rest>
__device__ __forceinline__ float2* tile_fft_static_rec(float2* a, float2* b, const float2* tw) {
    tile_stage_static<N, NL, NT, R, P, TW>(a, b, tw);
    if constexpr (sizeof...(Rest) == 0) return b;
    else return tile_fft_static_rec<N, NL, NT, P * R, TW + (R - 1) * P, Rest...>(b, a, tw);
}

template <int... Rs>
struct StaticPlan;

template <int NL, int... Rs>
struct StaticTile {
    using Plan = StaticPlan<Rs...>;
    static constexpr bool kStatic = true;
    static constexpr int kN = (Rs * ...);
    static constexpr int kNL = NL;
    static constexpr unsigned kMask = (radix_bit(Rs) | ...);
    // Forward FFT of the NL lines of a column tile held in a (ping-pong with
    // b) by an NT-thread block; ends synchronised and returns the buffer
    // holding the result.
    template <int NT>
    __device__ __forceinline__ static float2* tile_fft(float2* a, float2* b, const float2* tw) {
        return tile_fft_static_rec<kN, NL, NT, 1, 0, Rs...>(a, b, tw);
    }
};

// ---------------------------------------------------------------------------
// Compile-time slab FFTs: the in-place, register-resident stages of fft_lds
// (up to ceil(MAXV / R) butterflies per thread, read all, synchronise,
// write all) with R, P, the span and the twiddle offset as constants, in
// three layouts:
//   kLayTile:    element stride NL, line stride 1, consecutive threads on
//                consecutive lines (b = line + i NL);
//   kLayLines:   element stride 1, line stride N, consecutive threads on
//                consecutive butterflies of a line (b = i + line T);
//   kLayLinesLF: element stride 1, line stride N, consecutive threads on
//                consecutive lines (b = line + i nlines; nlines = NL when
//                NL > 0, else the runtime count with its float reciprocal).
// LDS holds 64 four-byte banks, a float2 two of them, so a half wave's 32
// float2 accesses are conflict-free when they fall on 32 distinct bank
// pairs.  In kLayLines a half wave that crosses from one line into the next
// reads element i of line l + 1 at pair (N + i) mod 32, which collides with
// the run of line l for every span T not congruent to N mod 32 (N = 175: T
// = 35 and 25 both collide).  In kLayLinesLF the lanes of one butterfly
// index sit N apart, and an odd N reaches 32 distinct pairs.
// ---------------------------------------------------------------------------

constexpr int kLayLines = 0, kLayTile = 1, kLayLinesLF = 2;

template <int N, int NT, int MAXV, int LAY, int NL, int R, int P, int TW>
__device__ __forceinline__ void slab_stage_static(float2* buf, const float2* __restrict__ tw, int nlines,
                                                  float nlinv) {
    constexpr bool TILE = LAY == kLayTile;
    constexpr int T = N / R;
    constexpr int MAXB = (MAXV + R - 1) / R;
    constexpr int ES = TILE ? NL : 1;
    const int nb = TILE ? T * NL : T * (LAY == kLayLinesLF && NL > 0 ? NL : nlines);
    float2 v[MAXB][R];
    int obase[MAXB];
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        const int b = (int)threadIdx.x + m * NT;
        obase[m] = -1;
        if (b < nb) {
            int i, ib;
            if constexpr (TILE) {
                i = b / NL;
                ib = b;
            } else if constexpr (LAY == kLayLinesLF) {
                int line;
                if constexpr (NL > 0) {
                    i = b / NL;
                    line = b - i * NL;
                } else {
                    i = fdiv(b, nlines, nlinv);
                    line = b - imul(i, nlines);
                }
                ib = line * N + i;
            } else {
                const int line = b / T;
                i = b - line * T;
                ib = line * N + i;
            }
#pragma unroll
            for (int j = 0; j < R; ++j) v[m][j] = buf[ib + j * T * ES];
            int k = 0;
            if constexpr (P > 1) {
                k = (int)((unsigned)i % (unsigned)P);
                const float2* w = tw + TW + k;
#pragma unroll
                for (int j = 1; j < R; ++j) v[m][j] = cmul(v[m][j], w[(j - 1) * P]);
            }
            dft<R>(v[m]);
            // (i - k) R + k = i + (R - 1)(i - k) elements along the line
            obase[m] = ib + (R - 1) * ES * (i - k);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        if (obase[m] >= 0) {
#pragma unroll
            for (int j = 0; j < R; ++j) buf[obase[m] + j * P * ES] = v[m][j];
        }
    }
    __syncthreads();
}

template <int N, int NT, int MAXV, int LAY, int NL, int P, int TW, int R, int... Rest>
__device__ __forceinline__ void slab_fft_static_rec(float2* buf, const float2* tw, int nlines, float nlinv) {
    slab_stage_static<N, NT, MAXV, LAY, NL, R, P, TW>(buf, tw, nlines, nlinv);
    if constexpr (sizeof...(Rest) > 0)
        slab_fft_static_rec<N, NT, MAXV, LAY, NL, P * R, TW + (R - 1) * P, Rest...>(buf, tw, nlines, nlinv);
}

template <int NL, int... Rs>
struct StaticTile;

// Stage radices in order, 4 bits each, stage 0 in the lowest bits (the
// descriptor's rcode).
template <int... Rs>
constexpr unsigned long long radix_code() {
    unsigned long long c = 0;
    int s = 0;
    ((c |= (unsigned long long)Rs << (4 * s++)), ...);
    return c;
}

template <int... Rs>
struct StaticPlan {
    static constexpr bool kStatic = true;
    static constexpr int kN = (Rs * ...);
    static constexpr unsigned kMask = (radix_bit(Rs) | ...);
    static constexpr unsigned long long kCode = radix_code<Rs...>();
    static constexpr int kNS = sizeof...(Rs);
    static constexpr int kR[kNS] = {Rs...};
    // product of the radices before stage s, and the float2 offset of stage
    // s's twiddles in the descriptor's table ([j - 1][k] per stage)
    static constexpr int P(int s) {
        int p = 1;
        for (int t = 0; t < s; ++t) p *= kR[t];
        return p;
    }
    static constexpr int TW(int s) {
        int o = 0, p = 1;
        for (int t = 0; t < s; ++t) {
            o += (kR[t] - 1) * p;
            p *= kR[t];
        }
        return o;
    }
    // the same radices as a column-tile plan of NL lines
    template <int NL>
    using Tile = StaticTile<NL, Rs...>;
    __host__ static bool same_radices(const int* r, int nr) {
        constexpr int want[] = {Rs...};
        if (nr != (int)sizeof...(Rs)) return false;
        for (int i = 0; i < nr; ++i)
            if (r[i] != want[i]) return false;
        return true;
    }
    // the N/2 + 1 half-spectrum columns of a slab, tile layout
    template <int NT, int MAXV>
    __device__ __forceinline__ static void tile_inplace(float2* buf, const float2* tw) {
        slab_fft_static_rec<kN, NT, MAXV, kLayTile, kN / 2 + 1, 1, 0, Rs...>(buf, tw, 0, 0.f);
    }
    // nlines contiguous lines of length N, consecutive threads on a line's
    // consecutive butterflies
    template <int NT, int MAXV>
    __device__ __forceinline__ static void lines_inplace(float2* buf, const float2* tw, int nlines) {
        slab_fft_static_rec<kN, NT, MAXV, kLayLines, 1, 1, 0, Rs...>(buf, tw, nlines, 0.f);
    }
    // the same lines with consecutive threads on consecutive lines; NLC > 0
    // fixes the line count at compile time, else nlines and its reciprocal
    template <int NT, int MAXV, int NLC = 0>
    __device__ __forceinline__ static void lines_inplace_lf(float2* buf, const float2* tw, int nlines) {
        slab_fft_static_rec<kN, NT, MAXV, kLayLinesLF, NLC, 1, 0, Rs...>(buf, tw, nlines,
                                                                          1.0f / (float)nlines);
    }
};

struct DynamicPlan {
    static constexpr bool kStatic = false;
    static constexpr int kN = 0;
};

// Stages S .. SEND - 1 of a compile-time plan, in place over a slab in
// layout LAY (the kernels that fuse a stage with their loads, stores or
// layout changes run the other stages through this).
template <class SP, int NT, int MAXV, int LAY, int NL, int S, int SEND>
__device__ __forceinline__ void slab_stages(float2* buf, const float2* tw, int nlines, float nlinv) {
    if constexpr (S < SEND) {
        slab_stage_static<SP::kN, NT, MAXV, LAY, NL, SP::kR[S], SP::P(S), SP::TW(S)>(buf, tw, nlines, nlinv);
        slab_stages<SP, NT, MAXV, LAY, NL, S + 1, SEND>(buf, tw, nlines, nlinv);
    }
}

// Stages S .. SEND - 1 of a compile-time plan over a column tile of NL
// lines (element stride NL), ping-pong from a to b; returns the buffer that
// holds the result (ends synchronised).
template <class PL, int NL, int NT, int S, int SEND>
__device__ __forceinline__ float2* tile_stages_pp(float2* a, float2* b, const float2* tw) {
    if constexpr (S >= SEND) {
        return a;
    } else {
        tile_stage_static<PL::kN, NL, NT, PL::kR[S], PL::P(S), PL::TW(S)>(a, b, tw);
        return tile_stages_pp<PL, NL, NT, S + 1, SEND>(b, a, tw);
    }
}

// The twiddles (k = i mod P) and radix-R DFT of one butterfly of a stage
// with P earlier radices and twiddles at offset TW: the arithmetic of
// slab_stage_static, for stages whose loads and stores are the caller's.
template <int R, int P, int TW>
__device__ __forceinline__ void twiddle_dft(float2 (&v)[R], int i, const float2* __restrict__ tw) {
    if constexpr (P > 1) {
        const int k = (int)((unsigned)i % (unsigned)P);
        const float2* w = tw + TW + k;
#pragma unroll
        for (int j = 1; j < R; ++j) v[j] = cmul(v[j], w[(j - 1) * P]);
    }
    dft<R>(v);
}

// fft_lds(buf, fd, nl, nl, 1, true) with nl = n / 2 + 1 columns
template <class SP, int MAXV, unsigned RM, int NT>
__device__ __forceinline__ void slab_fft_tile(float2* buf, const FftDesc& fd, int nl) {
    if constexpr (SP::kStatic) SP::template tile_inplace<NT, MAXV>(buf, fd.twid);
    else fft_lds<MAXV, RM>(buf, fd, nl, nl, 1, true);
}

// fft_lds(buf, fd, nlines, 1, n, false); a compile-time plan runs the lines
// with consecutive threads on consecutive lines (NLC > 0: nlines == NLC)
template <class SP, int MAXV, unsigned RM, int NT, int NLC = 0>
__device__ __forceinline__ void slab_fft_lines(float2* buf, const FftDesc& fd, int nlines) {
    if constexpr (SP::kStatic) SP::template lines_inplace_lf<NT, MAXV, NLC>(buf, fd.twid, nlines);
    else fft_lds<MAXV, RM>(buf, fd, nlines, 1, fd.n, false);
}

// The runtime-descriptor form (any length and tile width).
struct DynamicTile {
    static constexpr bool kStatic = false;
    static constexpr int kN = 0;
    static constexpr int kNL = 0;
    static constexpr unsigned kMask = 0;
};

template <class SP, unsigned RM, int NT>
__device__ __forceinline__ float2* tile_fft(float2* a, float2* b, const FftDesc& fd, int nl) {
    if constexpr (SP::kStatic) return SP::template tile_fft<NT>(a, b, fd.twid);
    else return fft_lds_pp<RM>(a, b, fd, nl, nl, 1, true);
}

}  // namespace vx

#pragma clang fp contract(off)
