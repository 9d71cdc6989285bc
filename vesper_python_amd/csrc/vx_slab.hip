// (y, z) slab kernels on compile-time FFT plans with their boundary passes
// folded into the FFT stages next to them: the forward slab reading its first
// stage from the query planes (k_fwd_zy_fused), the inverse slab reducing its
// last stage into the argmax (k_inv_yz_fused), and the same three folds in the
// pencil passes of the larger grids (k_fwd_z_fused, k_axis_y_fused,
// k_inv_z_fused; the z passes keep a chunk's lines interleaved in LDS, and
// the inverse one walks several chunks per workgroup with one argmax atomic).
// The separate-pass forms and the runtime-plan kernels live in
// vx_search.hip; these are the default on every grid edge with a
// compile-time plan (VX_TUNE_SLAB_FUSED).
// Their own translation unit, so that they compile beside the search
// kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <type_traits>

#include "vx_common.hpp"
#include "vx_fft.hpp"
#include "vx_internal.hpp"
#include "vx_plans.hpp"

namespace vx {

// The y half of a forward slab on a compile-time plan: lines [l0, l1) of
// the slab hold the z spectra of row pairs (2l, 2l + 1); the first y stage
// unpacks them (element (y, kz) is A or B of line y / 2, formed from Z[kz]
// and Z[n - kz]; rows outside the lines are zero without a read), and the
// last y stage, whose butterflies write the positions they read, stores the
// slab's [ky][kz] spectrum to out.
template <int MAXV, int NT, class SP>
__device__ __forceinline__ void fwd_slab_y(float2* lds, const float2* __restrict__ tw, int l0, int l1,
                                           float2* __restrict__ out) {
    constexpr int n = SP::kN;
    constexpr int nhz = n / 2 + 1;
    constexpr int NS = SP::kNS;
    // y stage 0 (P = 1) with the unpack Z = A + iB of line l = y / 2:
    // A[k] = (Z[k] + conj Z[n-k]) / 2, B[k] = (Z[k] - conj Z[n-k]) / 2i
    {
        constexpr int R = SP::kR[0], T = n / R, NB = T * nhz, MAXB = (MAXV + R - 1) / R;
        float2 v[MAXB][R];
        int ob[MAXB];
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            const int b = (int)threadIdx.x + m * NT;
            ob[m] = -1;
            if (b < NB) {
                const int i = b / nhz, col = b - i * nhz;
                const int cw = col == 0 ? 0 : n - col;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int y = i + j * T, l = y >> 1;
                    float2 a = make_float2(0.f, 0.f);
                    if (l >= l0 && l < l1) {
                        const float2 zk = lds[l * n + col];
                        const float2 wv = lds[l * n + cw];
                        a = (y & 1) ? make_float2(0.5f * (zk.y + wv.y), 0.5f * (wv.x - zk.x))
                                    : make_float2(0.5f * (zk.x + wv.x), 0.5f * (zk.y - wv.y));
                    }
                    v[m][j] = a;
                }
                dft<R>(v[m]);
                ob[m] = b + (R - 1) * nhz * i;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            if (ob[m] >= 0) {
#pragma unroll
                for (int j = 0; j < R; ++j) lds[ob[m] + j * nhz] = v[m][j];
            }
        }
        __syncthreads();
    }
    slab_stages<SP, NT, MAXV, kLayTile, nhz, 1, NS - 1>(lds, tw, 0, 0.f);
    // last y stage: butterfly (col, i < T = P) reads and writes rows i + j P
    {
        constexpr int S = NS - 1, R = SP::kR[S], P = SP::P(S), T = n / R, NB = T * nhz;
        constexpr int MAXB = (MAXV + R - 1) / R;
        static_assert(T == P, "the last stage of a plan spans its earlier radices");
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            const int b = (int)threadIdx.x + m * NT;
            if (b < NB) {
                const int i = b / nhz;
                float2 v[R];
#pragma unroll
                for (int j = 0; j < R; ++j) v[j] = lds[b + j * T * nhz];
                twiddle_dft<R, P, SP::TW(S)>(v, i, tw);
#pragma unroll
                for (int j = 0; j < R; ++j) out[b + j * P * nhz] = v[j];
            }
        }
    }
}

// k_fwd_zy on a compile-time plan with its boundary passes folded into the
// stages next to them (same arithmetic, same results):
//   * the first z stage reads its butterflies' inputs from the planes (two
//     real rows as one complex line; consecutive threads on consecutive
//     butterflies of a line, so the loads coalesce) instead of a load pass
//     through LDS;
//   * the half-spectrum unpack of each line pair happens inside the first y
//     stage's loads: element (y, kz) is A or B of line y / 2, formed from Z[kz]
//     and Z[n - kz], and rows outside the box's lines are zero without a read;
//   * the last y stage, whose butterflies write the positions they read,
//     stores to HBM instead of to LDS.
// This takes three LDS passes and six barriers out of the slab.
template <int MAXV, int NT, class SP>
__global__ void __launch_bounds__(NT)
k_fwd_zy_fused(const float* __restrict__ q, float2* __restrict__ spec, FftDesc fdg, int lo, int w) {
    static_assert(SP::kStatic && SP::kNS >= 2, "the fused slab needs a compile-time plan of two stages or more");
    constexpr int n = SP::kN;
    constexpr int nhz = n / 2 + 1;
    constexpr int NS = SP::kNS;
    extern __shared__ float2 lds[];
    const FftDesc fd = stage_twiddles(fdg, lds + n * nhz);
    const int x = lo + blockIdx.x;
    const size_t plane = blockIdx.y;
    const float* src = q + (plane * n + x) * (size_t)n * n;
    const int l0 = lo >> 1, l1 = ((lo + w - 1) >> 1) + 1;
    // z stage 0 (P = 1) from the planes; z outside [lo, lo + w) is zero
    {
        constexpr int R = SP::kR[0], T = n / R, MAXB = (MAXV + R - 1) / R;
        const int nb = T * (l1 - l0);
        float2 v[MAXB][R];
        int ob[MAXB];
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            const int b = (int)threadIdx.x + m * NT;
            ob[m] = -1;
            if (b < nb) {
                const int lr = b / T, i = b - lr * T, l = l0 + lr;
                const float* r0 = src + (size_t)(2 * l) * n;
                const bool hb = 2 * l + 1 < n;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int z = i + j * T;
                    float re = 0.f, im = 0.f;
                    if (z >= lo && z < lo + w) {
                        re = r0[z];
                        if (hb) im = r0[n + z];
                    }
                    v[m][j] = make_float2(re, im);
                }
                dft<R>(v[m]);
                ob[m] = l * n + i * R;
            }
        }
        // nothing has read the slab's LDS yet: write without a barrier
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            if (ob[m] >= 0) {
#pragma unroll
                for (int j = 0; j < R; ++j) lds[ob[m] + j] = v[m][j];
            }
        }
        __syncthreads();
    }
    slab_stages<SP, NT, MAXV, kLayLinesLF, 0, 1, NS>(lds + l0 * n, fd.twid, l1 - l0, 1.0f / (float)(l1 - l0));
    fwd_slab_y<MAXV, NT, SP>(lds, fd.twid, l0, l1, spec + (plane * n + x) * (size_t)n * nhz);
}

// k_inv_yz on a compile-time plan with its boundary passes folded into the
// stages next to them (same arithmetic, same results):
//   * the first y stage reads its butterflies' inputs (conjugated) from HBM,
//     coalesced along kz, instead of a load pass through LDS;
//   * the Hermitian pack of row pairs (2l, 2l + 1) into one complex line
//     happens inside the first z stage's loads (consecutive threads on
//     consecutive butterflies of a line, so the row reads stay on distinct
//     banks);
//   * the last z stage, whose butterflies write the positions they read,
//     reduces its outputs into the argmax (and writes the real maps when
//     asked) from registers instead of through LDS.
// Without real maps the last stage puts consecutive threads on consecutive
// lines (conflict-free LDS reads at odd n); with them, on consecutive
// butterflies of a line, so that the map stores coalesce along z.
template <int MAXV, int NT, class SP>
__global__ void __launch_bounds__(NT)
k_inv_yz_fused(const float2* __restrict__ sum, FftDesc fdg, int G, unsigned long long* __restrict__ keys,
               int nkeys, float* __restrict__ real_out) {
    static_assert(SP::kStatic && SP::kNS >= 2, "the fused slab needs a compile-time plan of two stages or more");
    constexpr int n = SP::kN;
    constexpr int nhz = n / 2 + 1;
    constexpr int npair = (n + 1) / 2;
    constexpr int half = n / 2;
    constexpr bool even = (n & 1) == 0;
    constexpr int NS = SP::kNS;
    extern __shared__ float2 lds[];
    const FftDesc fd = stage_twiddles(fdg, lds + n * nhz);
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    const int x = blockIdx.x;
    const int bg = blockIdx.y;
    const float2* src = sum + ((size_t)bg * n + x) * (size_t)n * nhz;
    // y stage 0 (P = 1) from HBM: the conj-forward transform of the y-inverse
    {
        constexpr int R = SP::kR[0], T = n / R, NB = T * nhz, MAXB = (MAXV + R - 1) / R;
        float2 v[MAXB][R];
        int ob[MAXB];
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            const int b = (int)threadIdx.x + m * NT;
            ob[m] = -1;
            if (b < NB) {
                const int i = b / nhz;
#pragma unroll
                for (int j = 0; j < R; ++j) v[m][j] = cconj(src[b + j * T * nhz]);
                dft<R>(v[m]);
                ob[m] = b + (R - 1) * nhz * i;
            }
        }
        // nothing has read the slab's LDS yet: write without a barrier
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            if (ob[m] >= 0) {
#pragma unroll
                for (int j = 0; j < R; ++j) lds[ob[m] + j * nhz] = v[m][j];
            }
        }
        __syncthreads();
    }
    slab_stages<SP, NT, MAXV, kLayTile, nhz, 1, NS>(lds, fd.twid, 0, 0.f);
    // z stage 0 (P = 1) with the pack: conj(Z), Z[k] = A[k] + i B[k] of rows
    // 2l, 2l + 1 (lds holds their conjugates); Hermitian extension for
    // k > n/2; imaginary parts of the DC (and Nyquist) bins dropped as c2r does
    {
        constexpr int R = SP::kR[0], T = n / R, NB = T * npair, MAXB = (MAXV + R - 1) / R;
        float2 v[MAXB][R];
        int ob[MAXB];
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            const int b = (int)threadIdx.x + m * NT;
            ob[m] = -1;
            if (b < NB) {
                const int l = b / T, i = b - l * T;
                const bool hasb = (2 * l + 1) < n;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int k = i + j * T;
                    if (k <= half) {
                        float2 ra = lds[(2 * l) * nhz + k];
                        float2 rb = hasb ? lds[(2 * l + 1) * nhz + k] : make_float2(0.f, 0.f);
                        if (k == 0 || (even && k == half)) { ra.y = 0.f; rb.y = 0.f; }
                        v[m][j] = make_float2(ra.x + rb.y, ra.y - rb.x);
                    } else {
                        const int kk = n - k;
                        const float2 ra = lds[(2 * l) * nhz + kk];
                        const float2 rb = hasb ? lds[(2 * l + 1) * nhz + kk] : make_float2(0.f, 0.f);
                        v[m][j] = make_float2(ra.x - rb.y, -ra.y - rb.x);
                    }
                }
                dft<R>(v[m]);
                ob[m] = l * n + i * R;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            if (ob[m] >= 0) {
#pragma unroll
                for (int j = 0; j < R; ++j) lds[ob[m] + j] = v[m][j];
            }
        }
        __syncthreads();
    }
    slab_stages<SP, NT, MAXV, kLayLinesLF, npair, 1, NS - 1>(lds, fd.twid, npair, 0.f);
    // last z stage + argmax: z[n] = conj(F), row 2l = F.x, row 2l+1 = -F.y
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    {
        constexpr int S = NS - 1, R = SP::kR[S], P = SP::P(S), T = n / R, NB = T * npair;
        constexpr int MAXB = (MAXV + R - 1) / R;
        static_assert(T == P, "the last stage of a plan spans its earlier radices");
        const size_t n3 = (size_t)n * n * n;
        const bool maps = real_out != nullptr;
#pragma unroll
        for (int m = 0; m < MAXB; ++m) {
            const int b = (int)threadIdx.x + m * NT;
            if (b < NB) {
                int l, i;
                if (maps) { l = b / T; i = b - l * T; }
                else      { i = b / npair; l = b - i * npair; }
                float2 v[R];
#pragma unroll
                for (int j = 0; j < R; ++j) v[j] = lds[l * n + i + j * T];
                twiddle_dft<R, P, SP::TW(S)>(v, i, fd.twid);
                const uint32_t row = ((uint32_t)x * n + 2 * l) * n;
                const bool hasb = 2 * l + 1 < n;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const uint32_t ia = row + i + j * P;
                    const float va = v[j].x;
                    if (better(va, ia, bv, bi)) { bv = va; bi = ia; }
                    if (maps) real_out[(size_t)bg * n3 + ia] = va;
                    if (hasb) {
                        const uint32_t ib = ia + n;
                        const float vb = -v[j].y;
                        if (better(vb, ib, bv, bi)) { bv = vb; bi = ib; }
                        if (maps) real_out[(size_t)bg * n3 + ib] = vb;
                    }
                }
            }
        }
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) {
        const int b = bg / G, g = bg - b * G;
        atomicMax(&keys[(size_t)b * nkeys + g], argmax_key(bv, bi));
    }
}

// ---------------------------------------------------------------------------
// Pencil passes (grids whose (y, z) slab exceeds LDS) on compile-time plans,
// with the same folding of boundary passes into the stages next to them.
// ---------------------------------------------------------------------------

// The line pairs of a pencil chunk (one per line of the z transform): the
// chunk's lines sit interleaved in LDS, element e of line l at e kPL + l
// (the tile layout), so that the lanes of a stage reach consecutive
// addresses whatever the line count (line-major lines of odd length N at
// 5..10 lines per chunk hit the same LDS banks).
template <class SP>
constexpr int pencil_lines() { return kPencilValues / SP::kN; }

// Forward r2c along z for a chunk of P line pairs of one x-slab (the
// k_fwd_z of vx_search.hip): the first stage reads its butterflies' two real
// rows from the planes (consecutive threads on consecutive butterflies of a
// line: coalesced, zero outside the box's z range without a load), the later
// stages run in place on the interleaved lines, then the unpack writes the
// rows' half spectra.  A last chunk of fewer than P pairs transforms
// leftover lines too and never reads them.
template <class SP>
__global__ void __launch_bounds__(256)
k_fwd_z_fused(const float* __restrict__ q, float2* __restrict__ spec, FftDesc fdg, int lo, int w, int P) {
    static_assert(SP::kStatic && SP::kNS >= 2, "the fused pencil needs a compile-time plan of two stages or more");
    constexpr int n = SP::kN;
    constexpr int nhz = n / 2 + 1;
    constexpr int NS = SP::kNS;
    constexpr int NT = 256;
    constexpr int PL = pencil_lines<SP>();   // == P (the launcher's check)
    extern __shared__ float2 lds[];
    float2* A = lds;
    const FftDesc fd = stage_twiddles(fdg, lds + P * n);
    const int x = lo + blockIdx.y;
    const size_t plane = blockIdx.z;
    const int l0 = (lo >> 1) + blockIdx.x * P;
    const int l1 = min(l0 + P, ((lo + w - 1) >> 1) + 1);
    const int np = l1 - l0;
    if (np <= 0) return;
    const float* src = q + (plane * n + x) * (size_t)n * n;
    {
        constexpr int R = SP::kR[0], T = n / R;
        const int nb = T * np;
#pragma unroll 1
        for (int b0 = 0; b0 < nb; b0 += NT) {
            const int b = b0 + (int)threadIdx.x;
            if (b < nb) {
                const int ll = b / T, i = b - ll * T, l = l0 + ll;
                const float* r0 = src + (size_t)(2 * l) * n;
                const bool hb = 2 * l + 1 < n;
                float2 v[R];
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int z = i + j * T;
                    float re = 0.f, im = 0.f;
                    if (z >= lo && z < lo + w) {
                        re = r0[z];
                        if (hb) im = r0[n + z];
                    }
                    v[j] = make_float2(re, im);
                }
                dft<R>(v);
                float2* O = A + i * R * PL + ll;   // nothing reads A before the barrier below
#pragma unroll
                for (int j = 0; j < R; ++j) O[j * PL] = v[j];
            }
        }
        __syncthreads();
    }
    slab_stages<SP, NT, kPencilMaxV, kLayTile, PL, 1, NS>(A, fd.twid, 0, 0.f);
    float2* out = spec + (plane * n + x) * (size_t)n * nhz;
    for (int u = threadIdx.x; u < np * nhz; u += NT) {
        const int ll = u / nhz, k = u - ll * nhz, l = l0 + ll;
        const float2 zk = A[k * PL + ll];
        const float2 wv = A[(k == 0 ? 0 : n - k) * PL + ll];
        out[(size_t)(2 * l) * nhz + k] = make_float2(0.5f * (zk.x + wv.x), 0.5f * (zk.y - wv.y));
        if (2 * l + 1 < n) out[(size_t)(2 * l + 1) * nhz + k] = make_float2(0.5f * (zk.y + wv.y), 0.5f * (wv.x - zk.x));
    }
}

// In-place complex FFT along y of a full tile of NL kz-columns of one x-slab
// (the k_axis_y of vx_search.hip on its compile-time tile plans): the first
// stage reads its inputs from HBM (rows outside [ylo, ylo + yw) as zero;
// conjugated for the inverse), the middle stages ping-pong in LDS, and the
// last stage, whose butterflies write the rows they read, stores to HBM.
// Workgroups reach the XCDs as in k_axis_y.
template <class TL>
__global__ void __launch_bounds__(256)
k_axis_y_fused(float2* __restrict__ buf, FftDesc fdg, int inverse, int ylo, int yw, int xlo, int ntile, int nxs,
               int total) {
    using PL = typename TL::Plan;
    static_assert(PL::kNS >= 2, "the fused y pass needs a compile-time plan of two stages or more");
    constexpr int n = PL::kN;
    constexpr int nhz = n / 2 + 1;
    constexpr int NL = TL::kNL;
    constexpr int NS = PL::kNS;
    constexpr int NT = 256;
    extern __shared__ float2 lds[];
    const int per = (total + 7) >> 3;
    const int m = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if ((int)(blockIdx.x >> 3) >= per || m >= total) return;
    const int tile = m % ntile, rest = m / ntile;
    const int xs = rest % nxs, plane = rest / nxs;
    float2* A = lds;
    float2* Bf = lds + n * NL;
    const FftDesc fd = stage_twiddles(fdg, lds + 2 * n * NL);
    float2* base = buf + ((size_t)plane * n + xlo + xs) * (size_t)n * nhz + tile * NL;
    {
        constexpr int R = PL::kR[0], T = n / R, NB = T * NL;
#pragma unroll 1
        for (int b0 = 0; b0 < NB; b0 += NT) {
            const int b = b0 + (int)threadIdx.x;
            if (NB % NT == 0 || b < NB) {
                const int i = b / NL, col = b - i * NL;
                float2 v[R];
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const int y = i + j * T;
                    float2 a = make_float2(0.f, 0.f);
                    if (y >= ylo && y < ylo + yw) a = base[(size_t)y * nhz + col];
                    v[j] = inverse ? cconj(a) : a;
                }
                dft<R>(v);
                float2* O = A + b + (R - 1) * NL * i;   // nothing reads A before the barrier below
#pragma unroll
                for (int j = 0; j < R; ++j) O[j * NL] = v[j];
            }
        }
        __syncthreads();
    }
    const float2* r = tile_stages_pp<PL, NL, NT, 1, NS - 1>(A, Bf, fd.twid);
    {
        constexpr int S = NS - 1, R = PL::kR[S], P = PL::P(S), T = n / R, NB = T * NL;
        static_assert(T == P, "the last stage of a plan spans its earlier radices");
#pragma unroll 1
        for (int b0 = 0; b0 < NB; b0 += NT) {
            const int b = b0 + (int)threadIdx.x;
            if (NB % NT == 0 || b < NB) {
                const int i = b / NL, col = b - i * NL;
                float2 v[R];
#pragma unroll
                for (int j = 0; j < R; ++j) v[j] = r[b + j * T * NL];
                twiddle_dft<R, P, PL::TW(S)>(v, i, fd.twid);
#pragma unroll
                for (int j = 0; j < R; ++j) base[(size_t)(i + j * P) * nhz + col] = inverse ? cconj(v[j]) : v[j];
            }
        }
    }
}

// Inverse c2r along z over chunks of P line pairs of one x-slab (the k_inv_z
// of vx_search.hip): the Hermitian pack of rows (2l, 2l + 1) of the
// y-inverted half spectrum happens inside the first stage's loads from HBM,
// the later stages run in place on the interleaved lines (pencil_lines),
// and the last stage reduces its outputs into the argmax (and writes the
// real maps when asked) from registers.  Each workgroup walks K chunks
// [blockIdx.x K, ...) and publishes one atomic: the keys of a launch share a
// cache line, where one atomic per chunk serialises (the same maximum either
// way: the key order is total).
template <class SP>
__global__ void __launch_bounds__(256)
k_inv_z_fused(const float2* __restrict__ sum, FftDesc fdg, int G, int P, int K, unsigned long long* __restrict__ keys,
              int nkeys, float* __restrict__ real_out) {
    static_assert(SP::kStatic && SP::kNS >= 2, "the fused pencil needs a compile-time plan of two stages or more");
    constexpr int n = SP::kN;
    constexpr int nhz = n / 2 + 1;
    constexpr int npair = (n + 1) / 2;
    constexpr int half = n / 2;
    constexpr bool even = (n & 1) == 0;
    constexpr int NS = SP::kNS;
    constexpr int NT = 256;
    constexpr int R0 = SP::kR[0], T0 = n / R0;
    constexpr int PL = pencil_lines<SP>();   // == P (the launcher's check)
    extern __shared__ float2 lds[];
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    float2* A = lds;
    const FftDesc fd = stage_twiddles(fdg, lds + P * n);
    const int x = blockIdx.y;
    const int bg = blockIdx.z;
    const int c0 = blockIdx.x * K;
    const int c1 = min(c0 + K, (npair + P - 1) / P);
    const float2* src = sum + ((size_t)bg * n + x) * (size_t)n * nhz;
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    const size_t n3 = (size_t)n * n * n;
    const bool maps = real_out != nullptr;
#pragma unroll 1
    for (int c = c0; c < c1; ++c) {
        const int l0 = c * P, np = min(P, npair - l0);
        {
            const int nb = T0 * np;
#pragma unroll 1
            for (int b0 = 0; b0 < nb; b0 += NT) {
                const int b = b0 + (int)threadIdx.x;
                if (b < nb) {
                    const int ll = b / T0, i = b - ll * T0, l = l0 + ll;
                    const bool hasb = (2 * l + 1) < n;
                    // every load of the butterfly first, then the pack
                    float2 ra[R0], rb[R0];
#pragma unroll
                    for (int j = 0; j < R0; ++j) {
                        const int k = i + j * T0;
                        const int kk = (k <= half) ? k : n - k;
                        ra[j] = src[(size_t)(2 * l) * nhz + kk];
                        rb[j] = hasb ? src[(size_t)(2 * l + 1) * nhz + kk] : make_float2(0.f, 0.f);
                    }
                    float2 v[R0];
#pragma unroll
                    for (int j = 0; j < R0; ++j) {
                        const int k = i + j * T0;
                        float2 a = ra[j], bb = rb[j];
                        if (k <= half) {
                            if (k == 0 || (even && k == half)) { a.y = 0.f; bb.y = 0.f; }
                            v[j] = make_float2(a.x - bb.y, -a.y - bb.x);       // conj(A + iB)
                        } else {
                            v[j] = make_float2(a.x + bb.y, a.y - bb.x);        // conj(conj A' + i conj B')
                        }
                    }
                    dft<R0>(v);
                    float2* O = A + i * R0 * PL + ll;   // the previous chunk's reads of A ended at a barrier
#pragma unroll
                    for (int j = 0; j < R0; ++j) O[j * PL] = v[j];
                }
            }
        }
        __syncthreads();
        slab_stages<SP, NT, kPencilMaxV, kLayTile, PL, 1, NS - 1>(A, fd.twid, 0, 0.f);
        {
            constexpr int S = NS - 1, R = SP::kR[S], P = SP::P(S), T = n / R;
            static_assert(T == P, "the last stage of a plan spans its earlier radices");
            // maps: consecutive threads along a line (coalesced stores);
            // else along the lines (consecutive LDS addresses)
            const int nb = maps ? T * np : T * PL;
#pragma unroll 1
            for (int b0 = 0; b0 < nb; b0 += NT) {
                const int b = b0 + (int)threadIdx.x;
                int ll, i;
                if (maps) { ll = b / T; i = b - ll * T; }
                else      { i = b / PL; ll = b - i * PL; }
                if (b < nb && ll < np) {
                    const int l = l0 + ll;
                    float2 v[R];
#pragma unroll
                    for (int j = 0; j < R; ++j) v[j] = A[(i + j * T) * PL + ll];
                    twiddle_dft<R, P, SP::TW(S)>(v, i, fd.twid);
                    const uint32_t row = ((uint32_t)x * n + 2 * l) * n;
                    const bool hasb = 2 * l + 1 < n;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const uint32_t ia = row + i + j * P;
                        const float va = v[j].x;
                        if (better(va, ia, bv, bi)) { bv = va; bi = ia; }
                        if (maps) real_out[(size_t)bg * n3 + ia] = va;
                        if (hasb) {
                            const uint32_t ib = ia + n;
                            const float vb = -v[j].y;
                            if (better(vb, ib, bv, bi)) { bv = vb; bi = ib; }
                            if (maps) real_out[(size_t)bg * n3 + ib] = vb;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) {
        const int b = bg / G, g = bg - b * G;
        atomicMax(&keys[(size_t)b * nkeys + g], argmax_key(bv, bi));
    }
}

#define VX_FUSED_LAUNCH(KER, MV, NTH, SPT, ...)                                        \
    do {                                                                               \
        allow_lds(KER<MV, NTH, SPT>, lds);                                             \
        hipLaunchKernelGGL((KER<MV, NTH, SPT>), grid, dim3(NTH), lds, st, __VA_ARGS__); \
    } while (0)

// hipGetLastError after a launch, as VX_LAUNCH_CHECK reports it.
static int launch_status() {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return VX_OK;
    set_error("kernel launch failed: %s (%s)", hipGetErrorString(e), __FILE__);
    return VX_ERR_HIP;
}

// k_fwd_zy_fused where the grid edge has a compile-time plan in its slab
// shape; returns false (nothing launched) otherwise.
bool launch_fwd_fused(const float* q, float2* spec, const FftDesc& fd, int n, int nplanes, int lo, int w,
                      hipStream_t st, int nt_pref, int* rc) {
    *rc = VX_OK;
    const int nhz = n / 2 + 1;
    int nthr, maxv;
    slab_shape(n, nhz, &nthr, &maxv, nt_pref);
    const size_t lds = ((size_t)n * nhz + fd.ntw) * sizeof(float2);
    if (maxv <= 0 || lds > kMaxLds) return false;
    dim3 grid(w, nplanes);
    bool hit = false;
    if (maxv == 8)
        hit = with_static_plan(fd, SmallLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_FUSED_LAUNCH(k_fwd_zy_fused, 8, 512, SP, q, spec, fd, lo, w);
        });
    else if (maxv == 17)
        hit = with_static_plan(fd, StaticLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_FUSED_LAUNCH(k_fwd_zy_fused, 16, 1024, SP, q, spec, fd, lo, w);
        });
    else if (maxv == 24)
        hit = with_static_plan(fd, MediumLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_FUSED_LAUNCH(k_fwd_zy_fused, 24, 1024, SP, q, spec, fd, lo, w);
        });
    if (hit) *rc = launch_status();
    return hit;
}

// k_inv_yz_fused where the grid edge has a compile-time plan in its slab
// shape; returns false (nothing launched) otherwise.
bool launch_inv_fused(const float2* sum, const FftDesc& fd, int n, int nbg, int G, unsigned long long* keys,
                      int nkeys, float* real_out, hipStream_t st, int nt_pref, int* rc) {
    *rc = VX_OK;
    const int nhz = n / 2 + 1;
    int nthr, maxv;
    slab_shape(n, nhz, &nthr, &maxv, nt_pref);
    const size_t lds = ((size_t)n * nhz + fd.ntw) * sizeof(float2);
    if (maxv <= 0 || lds > kMaxLds) return false;
    dim3 grid(n, nbg);
    bool hit = false;
    if (maxv == 8)
        hit = with_static_plan(fd, SmallLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_FUSED_LAUNCH(k_inv_yz_fused, 8, 512, SP, sum, fd, G, keys, nkeys, real_out);
        });
    else if (maxv == 17)
        hit = with_static_plan(fd, StaticLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_FUSED_LAUNCH(k_inv_yz_fused, 16, 1024, SP, sum, fd, G, keys, nkeys, real_out);
        });
    else if (maxv == 24)
        hit = with_static_plan(fd, MediumLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_FUSED_LAUNCH(k_inv_yz_fused, 24, 1024, SP, sum, fd, G, keys, nkeys, real_out);
        });
    if (hit) *rc = launch_status();
    return hit;
}

// The fused pencil passes where the grid edge has a compile-time plan
// (201..400); each returns false (nothing launched) otherwise.
bool launch_fwd_z_fused(const float* q, float2* spec, const FftDesc& fd, int n, int nplanes, int lo, int w,
                        hipStream_t st, int* rc) {
    *rc = VX_OK;
    const int P = pencil_pairs(n);
    // the kernels interleave pencil_lines = kPencilValues / n lines
    if (n * P > 256 * kPencilMaxV || P != kPencilValues / n) return false;
    const int l0 = lo >> 1, l1 = ((lo + w - 1) >> 1) + 1;
    const dim3 gz((l1 - l0 + P - 1) / P, w, nplanes);
    const size_t lds = ((size_t)P * n + fd.ntw) * sizeof(float2);
    if (lds > kMaxLds) return false;
    auto go = [&](auto sp) {
        using SP = decltype(sp);
        allow_lds(k_fwd_z_fused<SP>, lds);
        hipLaunchKernelGGL(k_fwd_z_fused<SP>, gz, dim3(256), lds, st, q, spec, fd, lo, w, P);
    };
    const bool hit = with_static_plan(fd, TileLengths{}, go) || with_static_plan(fd, LargeLengths{}, go);
    if (hit) *rc = launch_status();
    return hit;
}

bool launch_axis_y_fused(float2* buf, const FftDesc& fd, int n, int inverse, int ylo, int yw, int xlo, int nxs,
                         int nplanes, hipStream_t st, int* full_tiles, int* rc) {
    *rc = VX_OK;
    const int nhz = n / 2 + 1;
    const int T = pencil_T(n);
    const int full = nhz / T;
    *full_tiles = full;
    const size_t lds = (2 * (size_t)n * T + fd.ntw) * sizeof(float2);
    if (full <= 0 || lds > kMaxLds) return false;
    const int total = full * nxs * nplanes;
    const dim3 grid((unsigned)(((total + 7) / 8) * 8));
    auto go = [&](auto sp, auto width) {
        using TL = typename decltype(sp)::template Tile<decltype(width)::value>;
        allow_lds(k_axis_y_fused<TL>, lds);
        hipLaunchKernelGGL(k_axis_y_fused<TL>, grid, dim3(256), lds, st, buf, fd, inverse, ylo, yw, xlo, full, nxs,
                           total);
    };
    bool hit = false;
    if (T == 8) hit = with_static_plan(fd, TileLengths{}, [&](auto sp) { go(sp, std::integral_constant<int, 8>{}); });
    else if (T == 4) hit = with_static_plan(fd, LargeLengths{}, [&](auto sp) { go(sp, std::integral_constant<int, 4>{}); });
    if (hit) *rc = launch_status();
    return hit;
}

// Chunks of line pairs per workgroup of the inverse z pass, of nch per
// x-slab: `chunks` when positive, else equal shares that leave about 8192
// workgroups in the launch (32 per CU; one atomic and one twiddle fill per
// workgroup instead of one per chunk).
static int pencil_chunks(int nch, int n, int nbg, int chunks) {
    if (chunks <= 0) {
        const long long per = (long long)n * nbg;   // workgroups per chunk share
        const int shares = (int)std::min<long long>(nch, std::max<long long>(1, (8192 + per - 1) / per));
        chunks = (nch + shares - 1) / shares;
    }
    return chunks < 1 ? 1 : (chunks > nch ? nch : chunks);
}

bool launch_inv_z_fused(const float2* sum, const FftDesc& fd, int n, int nbg, int G, unsigned long long* keys,
                        int nkeys, float* real_out, hipStream_t st, int* rc, int chunks) {
    *rc = VX_OK;
    const int P = pencil_pairs(n);
    if (n * P > 256 * kPencilMaxV || P != kPencilValues / n) return false;
    const int npair = (n + 1) / 2;
    const int nch = (npair + P - 1) / P;
    const int K = pencil_chunks(nch, n, nbg, chunks);
    const dim3 gz((nch + K - 1) / K, n, nbg);
    const size_t lds = ((size_t)P * n + fd.ntw) * sizeof(float2);
    if (lds > kMaxLds) return false;
    auto go = [&](auto sp) {
        using SP = decltype(sp);
        allow_lds(k_inv_z_fused<SP>, lds);
        hipLaunchKernelGGL(k_inv_z_fused<SP>, gz, dim3(256), lds, st, sum, fd, G, P, K, keys, nkeys, real_out);
    };
    const bool hit = with_static_plan(fd, TileLengths{}, go) || with_static_plan(fd, LargeLengths{}, go);
    if (hit) *rc = launch_status();
    return hit;
}

}  // namespace vx
