// Compile-time FFT stage plans of the grid edges that have one, and the
// (y, z) slab workgroup shapes: shared by the search kernels (vx_search.hip)
// and the fused slab kernels (vx_slab.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "vx_common.hpp"
#include "vx_fft.hpp"
#include "vx_internal.hpp"

namespace vx {

// True when the descriptor's stages are exactly SP's radices in SP's order
// (the compile-time plan then reads the descriptor's twiddle table unchanged).
template <class SP>
inline bool plan_matches(const FftDesc& fd) {
    return fd.n == SP::kN && fd.rcode == SP::kCode;
}

// Grid edges with compile-time stage plans (radices in fft_desc_create's
// order): kernels instantiated for these lengths carry every stage's radix,
// span and twiddle offset as constants.  This list is every FFTW-fast length
// (2^a 3^b 5^c 7^d 11^e 13^f, e + f <= 1) from 90 to 180, the lengths whose
// (y, z) slab takes the 16 x 1024 workgroup shape; 175 = 5 * 5 * 7 is the
// grid edge of a 200^3 map at 2 A with the reference's padding.
template <class... Ps>
struct PlanList {};
using StaticLengths = PlanList<StaticPlan<2, 3, 3, 5>, StaticPlan<7, 13>, StaticPlan<8, 4, 3>,
    StaticPlan<2, 7, 7>, StaticPlan<3, 3, 11>, StaticPlan<4, 5, 5>, StaticPlan<8, 13>, StaticPlan<3, 5, 7>,
    StaticPlan<4, 3, 3, 3>, StaticPlan<2, 5, 11>, StaticPlan<8, 2, 7>, StaticPlan<3, 3, 13>,
    StaticPlan<8, 3, 5>, StaticPlan<5, 5, 5>, StaticPlan<2, 3, 3, 7>, StaticPlan<8, 8, 2>,
    StaticPlan<2, 5, 13>, StaticPlan<4, 3, 11>, StaticPlan<3, 3, 3, 5>, StaticPlan<4, 5, 7>,
    StaticPlan<8, 2, 3, 3>, StaticPlan<3, 7, 7>, StaticPlan<2, 3, 5, 5>, StaticPlan<2, 7, 11>,
    StaticPlan<4, 3, 13>, StaticPlan<8, 4, 5>, StaticPlan<2, 3, 3, 3, 3>, StaticPlan<3, 5, 11>,
    StaticPlan<8, 3, 7>, StaticPlan<5, 5, 7>, StaticPlan<8, 2, 11>, StaticPlan<4, 3, 3, 5>,
    StaticPlan<7, 5, 5>>;   // 175 with its radix-7 stage first (radix orders 1 and 3)

// The FFTW-fast lengths from 16 to 89 with two stages or more: their slab
// takes the 8 x 512 workgroup shape.
using SmallLengths = PlanList<StaticPlan<8, 2>, StaticPlan<2, 3, 3>, StaticPlan<4, 5>, StaticPlan<3, 7>,
    StaticPlan<2, 11>, StaticPlan<8, 3>, StaticPlan<5, 5>, StaticPlan<2, 13>, StaticPlan<3, 3, 3>,
    StaticPlan<4, 7>, StaticPlan<2, 3, 5>, StaticPlan<8, 4>, StaticPlan<3, 11>, StaticPlan<5, 7>,
    StaticPlan<4, 3, 3>, StaticPlan<3, 13>, StaticPlan<8, 5>, StaticPlan<2, 3, 7>, StaticPlan<4, 11>,
    StaticPlan<3, 3, 5>, StaticPlan<8, 2, 3>, StaticPlan<7, 7>, StaticPlan<2, 5, 5>, StaticPlan<4, 13>,
    StaticPlan<2, 3, 3, 3>, StaticPlan<5, 11>, StaticPlan<8, 7>, StaticPlan<4, 3, 5>, StaticPlan<3, 3, 7>,
    StaticPlan<8, 8>, StaticPlan<5, 13>, StaticPlan<2, 3, 11>, StaticPlan<2, 5, 7>, StaticPlan<8, 3, 3>,
    StaticPlan<3, 5, 5>, StaticPlan<7, 11>, StaticPlan<2, 3, 13>, StaticPlan<8, 2, 5>, StaticPlan<3, 3, 3, 3>,
    StaticPlan<4, 3, 7>, StaticPlan<8, 11>>;

// The FFTW-fast lengths from 181 to 200: their slab takes the 24 x 1024
// workgroup shape (above 200 the slab exceeds LDS and the pencils run).
using MediumLengths = PlanList<StaticPlan<2, 7, 13>, StaticPlan<3, 3, 3, 7>, StaticPlan<8, 8, 3>,
                               StaticPlan<3, 5, 13>, StaticPlan<4, 7, 7>, StaticPlan<2, 3, 3, 11>, StaticPlan<8, 5, 5>>;

// The FFTW-fast lengths from 201 to 256: the pencil passes carry their
// (y, z) transforms, and the register x pass (N x 8 <= 2048) its compile-time
// tile plan.
using TileLengths = PlanList<StaticPlan<8, 2, 13>, StaticPlan<2, 3, 5, 7>, StaticPlan<8, 3, 3, 3>,
    StaticPlan<4, 5, 11>, StaticPlan<8, 4, 7>, StaticPlan<3, 3, 5, 5>, StaticPlan<3, 7, 11>,
    StaticPlan<2, 3, 3, 13>, StaticPlan<8, 2, 3, 5>, StaticPlan<3, 3, 3, 3, 3>, StaticPlan<5, 7, 7>,
    StaticPlan<2, 5, 5, 5>, StaticPlan<4, 3, 3, 7>, StaticPlan<8, 8, 4>>;

// The FFTW-fast lengths from 257 to 400 (the larger grids of uneven map
// pairs): the register x pass takes them in 4-column tiles (N x 4 <= 2048)
// with a compile-time tile plan.
using LargeLengths = PlanList<StaticPlan<4, 5, 13>, StaticPlan<8, 3, 11>, StaticPlan<2, 3, 3, 3, 5>,
    StaticPlan<3, 7, 13>, StaticPlan<5, 5, 11>, StaticPlan<8, 5, 7>, StaticPlan<8, 4, 3, 3>, StaticPlan<2, 3, 7, 7>,
    StaticPlan<3, 3, 3, 11>, StaticPlan<4, 3, 5, 5>, StaticPlan<4, 7, 11>, StaticPlan<8, 3, 13>,
    StaticPlan<3, 3, 5, 7>, StaticPlan<8, 8, 5>, StaticPlan<4, 3, 3, 3, 3>, StaticPlan<5, 5, 13>,
    StaticPlan<2, 3, 5, 11>, StaticPlan<8, 2, 3, 7>, StaticPlan<7, 7, 7>, StaticPlan<2, 5, 5, 7>,
    StaticPlan<3, 3, 3, 13>, StaticPlan<8, 4, 11>, StaticPlan<8, 3, 3, 5>, StaticPlan<4, 7, 13>,
    StaticPlan<3, 5, 5, 5>, StaticPlan<2, 3, 3, 3, 7>, StaticPlan<8, 8, 2, 3>, StaticPlan<5, 7, 11>,
    StaticPlan<2, 3, 5, 13>, StaticPlan<8, 7, 7>, StaticPlan<4, 3, 3, 11>, StaticPlan<8, 2, 5, 5>>;

// Calls f(SP{}) for the listed plan whose stages are the descriptor's; false if none is.
template <class F, class... Ps>
inline bool with_static_plan(const FftDesc& fd, PlanList<Ps...>, F&& f) {
    bool hit = false;
    ((hit = hit || (plan_matches<Ps>(fd) && (f(Ps{}), true))), ...);
    return hit;
}

template <typename K>
inline int allow_lds(K kernel, size_t bytes) {
    if (bytes > 65536) VX_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return VX_OK;
}

// Slab kernels: choose threads and per-thread value budget so that every
// phase fits: z pass (npair x n), unpack (npair x nhz pairs, MAXV/2 each),
// y pass (nhz x n), inverse pack (npair x n).  Variants (MAXV, threads):
// (8, 512), (16, 512), (32, 512), (24, 1024).
// Lengths whose stages are radix 5 and 7 only (35, 175, 245, ...) get
// kernels compiled for just those radices: less code, fewer live scalars.
constexpr unsigned kRadix57 = radix_bit(5) | radix_bit(7);

inline bool slab_fits(int n, int nhz, int mv, int t) {
    const int npair = (n + 1) / 2;
    return n * nhz <= t * mv && npair * n <= t * mv && npair * nhz <= t * ((mv + 1) / 2);
}

// Slab workgroup shape: values per thread (MAXV) and threads.  nt_pref 0 or
// 1024 takes the 16 x 1024 shape (code 17) where it fits and the slab needs
// more than 8 x 512 (faster on MI355X: twice the waves per slab); 512 keeps
// the 512-thread shapes.
inline void slab_shape(int n, int nhz, int* nthr, int* maxv, int nt_pref = 0) {
    if (nt_pref != 512 && slab_fits(n, nhz, 16, 1024) && !slab_fits(n, nhz, 8, 512)) {
        *nthr = 1024; *maxv = 17; return;
    }
    const int npair = (n + 1) / 2;
    const int opts[4][2] = {{8, 512}, {16, 512}, {32, 512}, {24, 1024}};
    for (auto& o : opts) {
        const int mv = o[0], t = o[1];
        if (n * nhz <= t * mv && npair * n <= t * mv && npair * nhz <= t * ((mv + 1) / 2)) {
            *nthr = t; *maxv = mv; return;
        }
    }
    *nthr = 1024; *maxv = 0;
}

// Values per thread of a pencil workgroup's in-place z pass on a
// compile-time plan.
constexpr int kPencilMaxV = 8;

// Values per pencil workgroup: ~2k complex (32 KB of ping-pong LDS, four to
// five 256-thread workgroups per CU).  At ~4k (two workgroups per CU) the
// pencil searches ran 8-39 % slower (N = 208..343): too few waves to hide the
// LDS latency of the stages; at ~1k the shorter lines per workgroup lost more
// than the extra waves gained.
constexpr int kPencilValues = 2048;

inline int pencil_pairs(int n) {
    int P = kPencilValues / n;
    if (P < 1) P = 1;
    if (P > 32) P = 32;
    return P;
}

inline int pencil_T(int n) {
    // columns per y-pass workgroup: a power of two (runs of 8 T bytes stay
    // aligned to segments), at most 16
    int T = 16;
    while (T > 1 && T * n > kPencilValues) T >>= 1;
    return T;
}

}  // namespace vx
