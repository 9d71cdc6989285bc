// Rotational search for one map pair on MI355X: rotate-and-gather (b), 3-D
// FFT cross-correlation with fused conjugate multiply, channel sum and
// translation argmax (c), per batch of rotations.
//
// Pipeline per batch of B rotations (grid edge N, NHZ = N/2 + 1):
//   k_gather        query channels q[b][c][x][y][z] (float32), rotated and
//                   gathered as rot_mrc / rot_mrc_prob do
//                   (VESPER_1_prob_new.py:279-354, :1211-1299)
//   k_fwd_zy        per x-slab in LDS: r2c along z (two real lines packed
//                   into one complex line), complex FFT along y
//   k_xcorr         per tile of (ky, kz) columns: FFT along x, times the
//                   conjugated target spectrum, summed over the channels of a
//                   group, inverse FFT along x
//   k_inv_yz        per x-slab in LDS: inverse FFT along y, c2r along z
//                   (pair packing again), max + first argmax in C order
//                   -> one 64-bit atomicMax per slab
//   k_finalize      keys -> (best float32, translation)
// The channel sum happens in the frequency domain (linearity of the inverse
// transform), so there is one inverse per group instead of one per channel
// (VESPER_1_prob_new.py:464-491 + :357-374 compute C inverses then sum).
// On the grid edges with compile-time FFT plans the slab kernels (and, above
// N = 200, the pencil passes that replace them) run in the folded forms of
// vx_slab.hip: first stage read from HBM, unpack / pack inside the next
// stage, last stage stored or reduced into the argmax from registers.  The
// kernels here are the separate-pass and runtime-plan forms.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include <new>
#include <type_traits>

#include "vx_common.hpp"
#include "vx_fft.hpp"
#include "vx_internal.hpp"
#include "vx_gather.hpp"
#include "vx_plans.hpp"

namespace vx {

// ---------------------------------------------------------------------------
// FFT descriptor construction
// ---------------------------------------------------------------------------

// Stage radices of a length n: powers of two in radix 8, 4, 2, then the odd
// primes.  `order` permutes the stages (the autotuner's radix-order knob):
// 0 = that order; 1 = reversed; 2 = the odd primes first (ascending), then the
// powers of two; 3 = the odd primes first, descending, then the powers of two.
static bool factorize(int n, std::vector<int>& r, int order = 0) {
    r.clear();
    int m = n;
    std::vector<int> two, odd;
    while (m % 8 == 0) { two.push_back(8); m /= 8; }
    while (m % 4 == 0) { two.push_back(4); m /= 4; }
    while (m % 2 == 0) { two.push_back(2); m /= 2; }
    const int primes[] = {3, 5, 7, 11, 13};
    for (int p : primes)
        while (m % p == 0) { odd.push_back(p); m /= p; }
    if (order == 3) std::reverse(odd.begin(), odd.end());
    if (order >= 2) {
        r = odd;
        r.insert(r.end(), two.begin(), two.end());
    } else {
        r = two;
        r.insert(r.end(), odd.begin(), odd.end());
        if (order == 1) std::reverse(r.begin(), r.end());
    }
    return m == 1 && (int)r.size() <= kMaxStages;
}

int fft_desc_create(int n, FftDesc* d, float2** twid_out, int order) {
    std::vector<int> r;
    VX_REQUIRE(n >= 2 && factorize(n, r, order), VX_ERR_UNSUPPORTED,
               "grid edge %d is not 2^a 3^b 5^c 7^d 11^e 13^f", n);
    memset(d, 0, sizeof(*d));
    d->n = n;
    d->nst = (int)r.size();
    for (int s = 0; s < d->nst; ++s) d->rcode |= (unsigned long long)r[s] << (4 * s);
    std::vector<FftStage> stages(d->nst);
    std::vector<float2> tw;
    int p = 1;
    for (int s = 0; s < d->nst; ++s) {
        const int R = r[s];
        stages[s].R = R;
        stages[s].P = p;
        stages[s].Pinv = 1.0f / (float)p;
        stages[s].tw = (int)tw.size();
        d->radix_mask |= radix_bit(R);
        // [j - 1][k]: lanes holding consecutive k read consecutive twiddles
        for (int j = 1; j < R; ++j)
            for (int k = 0; k < p; ++k) {
                const double a = -2.0 * M_PI * (double)j * (double)k / ((double)p * (double)R);
                tw.push_back(make_float2((float)cos(a), (float)sin(a)));
            }
        p *= R;
    }
    if (tw.empty()) tw.push_back(make_float2(1.f, 0.f));
    d->ntw = (int)tw.size();
    // one allocation: stage table (padded to 256 B) then twiddles
    const size_t sbytes = 256;
    char* dev = nullptr;
    VX_HIP(hipMalloc(&dev, sbytes + tw.size() * sizeof(float2)));
    VX_HIP(hipMemcpy(dev, stages.data(), stages.size() * sizeof(FftStage), hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(dev + sbytes, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    d->st = (const FftStage*)dev;
    d->twid = (const float2*)(dev + sbytes);
    *twid_out = (float2*)dev;
    return VX_OK;
}

// ---------------------------------------------------------------------------
// Rotate-and-gather
// ---------------------------------------------------------------------------

template <int KIND>
__global__ void __launch_bounds__(256)
k_gather(const float4* __restrict__ svec, const float2* __restrict__ sscal,
         const float4* __restrict__ sprob, const double* __restrict__ sdat,
         const float* __restrict__ rot, int n, int nzb, int fma,
         float* __restrict__ q, double* __restrict__ qd,
         unsigned long long* __restrict__ keys, int nkeys, int lo, int w, int slo, int sw) {
    const int b = blockIdx.z;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < nkeys) keys[(size_t)b * nkeys + threadIdx.x] = 0ull;
    // destination box [lo, lo + w)^3 holds every voxel that can receive a
    // non-zero value; the planes are zero outside it
    const int x = lo + blockIdx.y;
    const int y = lo + (blockIdx.x / nzb) * 4 + (threadIdx.x >> 6);
    const int z = lo + (blockIdx.x % nzb) * 64 + (threadIdx.x & 63);
    if (y >= lo + w || z >= lo + w) return;
    float F[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = rot[b * 9 + i];
    constexpr int NOUT = (KIND == KIND_VEC) ? 3 : (KIND == KIND_PROB ? 7 : 1);
    float out[NOUT];
    double outd;
    gather_voxel<KIND>(x, y, z, n, slo, sw, F, fma != 0, svec, sscal, sprob, sdat, out, &outd);
    const size_t n3 = (size_t)n * n * n;
    const size_t v = ((size_t)x * n + y) * n + z;
    if (KIND == KIND_DATA64) {
        qd[(size_t)b * n3 + v] = outd;
    } else {
#pragma unroll
        for (int c = 0; c < NOUT; ++c) q[((size_t)b * NOUT + c) * n3 + v] = out[c];
    }
}

// Laplacian of gathered float64 data -> float32 query channel
// (laplacian_filter, VESPER_1_prob_new.py:494-526, then np.copyto to float32).
__global__ void k_lap_to_q(const double* __restrict__ qd, float* __restrict__ q, int n) {
    const size_t n3 = (size_t)n * n * n;
    const int b = blockIdx.y;
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n3) return;
    const double* a = qd + (size_t)b * n3;
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
    q[(size_t)b * n3 + v] = (float)r;
}

// ---------------------------------------------------------------------------
// Source preparation
// ---------------------------------------------------------------------------

__global__ void k_build_src(const double* __restrict__ data, const float* __restrict__ vec,
                            const float* __restrict__ prob, size_t n3, int mode, double ave,
                            float4* __restrict__ svec, float2* __restrict__ sscal,
                            float4* __restrict__ sprob, int n, int lo, int w) {
    // one thread per voxel of the box [lo, lo + w)^3: j indexes the compact
    // arrays, i the grid
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (size_t)w * w * w) return;
    const int bz = (int)(j % w), by = (int)((j / w) % w), bx = (int)(j / ((size_t)w * w));
    const size_t i = ((size_t)(lo + bx) * n + (lo + by)) * n + (lo + bz);
    const double d = data[i];
    const float mask = (d != 0.0) ? 1.f : 0.f;
    if (mode == VX_MODE_VEC || mode == VX_MODE_PROB) {
        svec[j] = make_float4(vec[3 * i], vec[3 * i + 1], vec[3 * i + 2], mask);
        if (mode == VX_MODE_PROB)
            sprob[j] = make_float4(prob[i], prob[n3 + i], prob[2 * n3 + i], prob[3 * n3 + i]);
    } else {
        // query channel value per mode (VESPER_1_prob_new.py:548-553), cast
        // to float32 as np.copyto into the float32 FFT input does
        float v = 0.f;
        if (d > 0) {
            if (mode == VX_MODE_OVERLAP) v = 1.f;
            else if (mode == VX_MODE_CC) v = (float)d;
            else if (mode == VX_MODE_PCC) v = (float)(d - ave);
        }
        sscal[j] = make_float2(v, mask);
    }
}

// Target channels (VESPER_1_prob_new.py:601-615, :831-838) as float32 planes.
__global__ void k_target_chan(const double* __restrict__ data, const float* __restrict__ vec,
                              const float* __restrict__ prob, const double* __restrict__ lap,
                              size_t n3, int mode, double ave, float* __restrict__ q) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    if (mode == VX_MODE_VEC || mode == VX_MODE_PROB) {
        q[i] = vec[3 * i];
        q[n3 + i] = vec[3 * i + 1];
        q[2 * n3 + i] = vec[3 * i + 2];
        if (mode == VX_MODE_PROB)
            for (int k = 0; k < 4; ++k) q[(3 + k) * n3 + i] = prob[k * n3 + i];
        return;
    }
    const double d = data[i];
    float v = 0.f;
    if (mode == VX_MODE_LAPLACIAN) v = (float)lap[i];
    else if (d > 0) {
        if (mode == VX_MODE_OVERLAP) v = 1.f;
        else if (mode == VX_MODE_CC) v = (float)d;
        else if (mode == VX_MODE_PCC) v = (float)(d - ave);
    }
    q[i] = v;
}

// ---------------------------------------------------------------------------
// FFT passes
// ---------------------------------------------------------------------------

// Forward: real slab q[plane][x] ([y][z]) -> spec[plane][x] ([ky][kz]).
template <int MAXV, int NT, unsigned RM = 0xffffffffu, class SP = DynamicPlan>
__global__ void __launch_bounds__(NT)
k_fwd_zy(const float* __restrict__ q, float2* __restrict__ spec, FftDesc fdg, int n_, int nhz_, int lo, int w) {
    // slabs x in [lo, lo + w) only; rows y outside [lo, lo + w) are zero, so
    // only the line pairs [l0, l1) covering them are transformed along z
    // (a compile-time plan fixes the grid edge, and with it the index math)
    const int n = SP::kStatic ? SP::kN : n_;
    const int nhz = SP::kStatic ? SP::kN / 2 + 1 : nhz_;
    extern __shared__ float2 lds[];
    const FftDesc fd = stage_twiddles(fdg, lds + n * nhz);
    const int x = lo + blockIdx.x;
    const size_t plane = blockIdx.y;
    const float* src = q + (plane * n + x) * (size_t)n * n;
    const int npair = (n + 1) >> 1;
    const int l0 = lo >> 1, l1 = ((lo + w - 1) >> 1) + 1;
    const int nact = (l1 - l0) * n;
    const int nthr = blockDim.x;
    {
        // all loads of the slab in flight before the first LDS write
        const float ninv = 1.0f / (float)n;
        float re[MAXV], im[MAXV];
#pragma unroll
        for (int m = 0; m < MAXV; ++m) {
            const int e = threadIdx.x + m * nthr;
            re[m] = 0.f;
            im[m] = 0.f;
            if (e < nact) {
                const int l = l0 + qdiv<SP::kStatic>(e, n, ninv), z = e - (l - l0) * n;
                if (z >= lo && z < lo + w) {   // the planes are zero outside the box: no load
                    re[m] = src[(size_t)(2 * l) * n + z];
                    if (2 * l + 1 < n) im[m] = src[(size_t)(2 * l + 1) * n + z];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < MAXV; ++m) {
            const int e = threadIdx.x + m * nthr;
            if (e < nact) lds[l0 * n + e] = make_float2(re[m], im[m]);
        }
    }
    __syncthreads();
    slab_fft_lines<SP, MAXV, RM, NT>(lds + l0 * n, fd, l1 - l0);
    // unpack Z = A + iB: A[k] = (Z[k] + conj Z[n-k]) / 2, B[k] = (Z[k] - conj Z[n-k]) / 2i
    {
        constexpr int MU = (MAXV + 1) / 2;
        float2 A[MU], B[MU];
        int dst[MU];
        const int nu = npair * nhz;
        const float nhzinv = 1.0f / (float)nhz;
#pragma unroll
        for (int m = 0; m < MU; ++m) {
            const int u = threadIdx.x + m * nthr;
            dst[m] = -1;
            if (u < nu) {
                const int l = qdiv<SP::kStatic>(u, nhz, nhzinv), k = u - l * nhz;
                A[m] = make_float2(0.f, 0.f);
                B[m] = make_float2(0.f, 0.f);
                if (l >= l0 && l < l1) {
                    const float2 zk = lds[l * n + k];
                    const float2 wv = lds[l * n + (k == 0 ? 0 : n - k)];
                    A[m] = make_float2(0.5f * (zk.x + wv.x), 0.5f * (zk.y - wv.y));
                    B[m] = make_float2(0.5f * (zk.y + wv.y), 0.5f * (wv.x - zk.x));
                }
                dst[m] = (2 * l) * nhz + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < MU; ++m) {
            if (dst[m] >= 0) {
                lds[dst[m]] = A[m];
                if (dst[m] + nhz < n * nhz) lds[dst[m] + nhz] = B[m];
            }
        }
        __syncthreads();
    }
    slab_fft_tile<SP, MAXV, RM, NT>(lds, fd, nhz);
    float2* out = spec + (plane * n + x) * (size_t)n * nhz;
#pragma unroll
    for (int m = 0; m < MAXV; ++m) {
        const int e = threadIdx.x + m * nthr;
        if (e < n * nhz) out[e] = lds[e];
    }
}

// x-axis: per tile of T flattened (ky, kz) columns.  For each channel c:
// FFT along x, multiply by tgt[c] (already conjugated), accumulate into its
// group; then inverse FFT along x of each group.
//   MODE_FWD_CONJ (tgt == nullptr): forward only, write conj(result) per
//   channel (target preparation).
template <int MAXV, int G>
__global__ void __launch_bounds__(256)
k_xcorr(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
        FftDesc fdg, int n, int nhz, int C, int T, unsigned grp_mask) {
    extern __shared__ float2 lds[];
    const FftDesc fd = stage_twiddles(fdg, lds + n * T);
    const int plane = n * nhz;
    const int j0 = blockIdx.x * T;
    const int nt = min(T, plane - j0);
    const int b = blockIdx.y;
    const int nthr = blockDim.x;
    const int ne = n * nt;
    float2 acc[G][MAXV];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int m = 0; m < MAXV; ++m) acc[g][m] = make_float2(0.f, 0.f);
    for (int c = 0; c < C; ++c) {
        const float2* src = spec + ((size_t)b * C + c) * n * (size_t)plane;
        for (int e = threadIdx.x; e < ne; e += nthr) {
            const int xx = e / nt, t = e - xx * nt;
            lds[e] = src[(size_t)xx * plane + j0 + t];
        }
        __syncthreads();
        fft_lds<MAXV>(lds, fd, nt, nt, 1, true);
        if (tgt == nullptr) {
            float2* dst = out + ((size_t)b * C + c) * n * (size_t)plane;
            for (int e = threadIdx.x; e < ne; e += nthr) {
                const int xx = e / nt, t = e - xx * nt;
                dst[(size_t)xx * plane + j0 + t] = cconj(lds[e]);
            }
            __syncthreads();
            continue;
        }
        const float2* tc = tgt + (size_t)c * n * plane;
        const int g = (grp_mask >> c) & 1;
#pragma unroll
        for (int m = 0; m < MAXV; ++m) {
            const int e = threadIdx.x + m * nthr;
            if (e < ne) {
                const int xx = e / nt, t = e - xx * nt;
                const float2 p = cmul(lds[e], tc[(size_t)xx * plane + j0 + t]);
#pragma unroll
                for (int gg = 0; gg < G; ++gg)
                    if (gg == g) acc[gg][m] = cadd(acc[gg][m], p);
            }
        }
        __syncthreads();
    }
    if (tgt == nullptr) return;
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int m = 0; m < MAXV; ++m) {
            const int e = threadIdx.x + m * nthr;
            if (e < ne) lds[e] = cconj(acc[g][m]);
        }
        __syncthreads();
        fft_lds<MAXV>(lds, fd, nt, nt, 1, true);
        float2* dst = out + ((size_t)b * G + g) * n * (size_t)plane;
        for (int e = threadIdx.x; e < ne; e += nthr) {
            const int xx = e / nt, t = e - xx * nt;
            dst[(size_t)xx * plane + j0 + t] = cconj(lds[e]);
        }
        __syncthreads();
    }
}

// x pass, correlate mode: one workgroup per (tile of T flattened (ky, kz)
// columns, rotation b).  For each channel: FFT along x in two ping-pong LDS
// buffers, times the conjugated target tile, accumulated per group in
// registers; then the inverse FFT along x of each group.  Block ids are
// laid out so that the B workgroups of one tile share an XCD (id mod 8) and
// are dispatched together: the target tile is read from HBM once per batch
// and from that XCD's L2 by the other rotations.

template <int MAXE, int G>
__global__ void __launch_bounds__(256)
k_xcorr_pp(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
           FftDesc fdg, int n, int nhz, int nb, int C, int T, int ntiles, unsigned grp_mask, int lo, int w) {
    extern __shared__ float2 lds[];
    float2* A = lds;
    float2* Bf = lds + n * T;
    const FftDesc fd = stage_twiddles(fdg, lds + 2 * n * T);
    const int id = blockIdx.x;
    int b, tile;
    if (!xcd_tile(id, nb, ntiles, &tile, &b)) return;
    const int plane = n * nhz;
    const int j0 = tile * T;
    const int nt = min(T, plane - j0);
    const int ne = n * nt;
    const float ntinv = 1.0f / (float)nt;
    float2 acc[G][MAXE];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int m = 0; m < MAXE; ++m) acc[g][m] = make_float2(0.f, 0.f);
    // software pipeline: the query tile of channel c + 1 and the target tile
    // of channel c are in flight while the FFT of channel c runs in LDS
    // (the workgroup barrier waits only for LDS traffic)
    float2 v[MAXE], tv[MAXE];
    // element e = threadIdx.x + 256 m of the tile: global offset, or -1 past
    // the tile; bit 30 marks a slab outside the query support
    int off[MAXE];
#pragma unroll
    for (int m = 0; m < MAXE; ++m) {
        const int e = threadIdx.x + m * 256;
        off[m] = -1;
        if (e < ne) {
            const int xx = fdiv(e, nt, ntinv), t = e - xx * nt;
            off[m] = xx * plane + t;
            if (xx < lo || xx >= lo + w) off[m] |= (1 << 30);
        }
    }
    auto load_query = [&](int c) {
        const float2* src = spec + ((size_t)b * C + c) * n * (size_t)plane + j0;
#pragma unroll
        for (int m = 0; m < MAXE; ++m) {
            if (off[m] >= 0)
                v[m] = (off[m] & (1 << 30)) ? make_float2(0.f, 0.f) : src[off[m]];
        }
    };
    load_query(0);
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int m = 0; m < MAXE; ++m) {
            const int e = threadIdx.x + m * 256;
            if (e < ne) A[e] = v[m];
        }
        __syncthreads();
        {
            const float2* tc = tgt + (size_t)c * n * plane + j0;
#pragma unroll
            for (int m = 0; m < MAXE; ++m)
                if (off[m] >= 0) tv[m] = tc[off[m] & ~(1 << 30)];
        }
        if (c + 1 < C) load_query(c + 1);
        const float2* r = fft_lds_pp(A, Bf, fd, nt, nt, 1, true);
        const int g = (grp_mask >> c) & 1;
#pragma unroll
        for (int m = 0; m < MAXE; ++m) {
            const int e = threadIdx.x + m * 256;
            if (e < ne) {
                const float2 pr = cmul(r[e], tv[m]);
#pragma unroll
                for (int gg = 0; gg < G; ++gg)
                    if (gg == g) acc[gg][m] = cadd(acc[gg][m], pr);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int m = 0; m < MAXE; ++m) {
            const int e = threadIdx.x + m * 256;
            if (e < ne) A[e] = cconj(acc[g][m]);
        }
        __syncthreads();
        const float2* r = fft_lds_pp(A, Bf, fd, nt, nt, 1, true);
        float2* dst = out + ((size_t)b * G + g) * n * (size_t)plane + j0;
#pragma unroll
        for (int m = 0; m < MAXE; ++m) {
            const int e = threadIdx.x + m * 256;
            if (off[m] >= 0) dst[off[m] & ~(1 << 30)] = cconj(r[e]);
        }
        __syncthreads();
    }
}

// x pass, LDS-accumulating variant: same work as k_xcorr_pp, but the group
// sums live in LDS (one more tile buffer) instead of registers, so a thread
// holds a single butterfly; smaller tiles give several workgroups per CU.
template <int G>
__global__ void __launch_bounds__(256)
k_xcorr_lds(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
            FftDesc fdg, int n, int nhz, int nb, int C, int T, int ntiles, unsigned grp_mask, int lo, int w) {
    extern __shared__ float2 lds[];
    const int id = blockIdx.x;
    int b, tile;
    if (!xcd_tile(id, nb, ntiles, &tile, &b)) return;
    float2* A = lds;
    float2* Bf = lds + n * T;
    float2* acc = lds + 2 * n * T;                 // [G][n * T]
    const FftDesc fd = stage_twiddles(fdg, lds + (2 + G) * n * T);
    const int plane = n * nhz;
    const int j0 = tile * T;
    const int nt = min(T, plane - j0);
    const int ne = n * nt;
    const float ntinv = 1.0f / (float)nt;
    for (int e = threadIdx.x; e < G * ne; e += 256) acc[e] = make_float2(0.f, 0.f);
    for (int c = 0; c < C; ++c) {
        const float2* src = spec + ((size_t)b * C + c) * n * (size_t)plane + j0;
        for (int e0 = threadIdx.x; e0 < ne; e0 += 4 * 256) {
            float2 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * 256;
                v[u] = make_float2(0.f, 0.f);
                if (e < ne) {
                    const int xx = fdiv(e, nt, ntinv), t = e - xx * nt;
                    if (xx >= lo && xx < lo + w) v[u] = src[(size_t)xx * plane + t];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (e0 + u * 256 < ne) A[e0 + u * 256] = v[u];
        }
        __syncthreads();
        const float2* r = fft_lds_pp(A, Bf, fd, nt, nt, 1, true);
        const float2* tc = tgt + (size_t)c * n * plane + j0;
        float2* ag = acc + ((grp_mask >> c) & 1) * ne;
        for (int e0 = threadIdx.x; e0 < ne; e0 += 4 * 256) {
            float2 tv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * 256;
                if (e < ne) {
                    const int xx = fdiv(e, nt, ntinv), t = e - xx * nt;
                    tv[u] = tc[(size_t)xx * plane + t];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * 256;
                if (e < ne) ag[e] = cadd(ag[e], cmul(r[e], tv[u]));
            }
        }
        __syncthreads();
    }
    for (int g = 0; g < G; ++g) {
        for (int e = threadIdx.x; e < ne; e += 256) A[e] = cconj(acc[g * ne + e]);
        __syncthreads();
        const float2* r = fft_lds_pp(A, Bf, fd, nt, nt, 1, true);
        float2* dst = out + ((size_t)b * G + g) * n * (size_t)plane + j0;
        for (int e = threadIdx.x; e < ne; e += 256) {
            const int xx = fdiv(e, nt, ntinv), t = e - xx * nt;
            dst[(size_t)xx * plane + t] = cconj(r[e]);
        }
        __syncthreads();
    }
}

// LDS-accumulating x pass with software prefetch (tiles of n * T <= 256 * PF
// values): the next channel's spectra and this channel's target tile are
// loaded into registers before the FFT starts, so their HBM latency overlaps
// the transform (barriers wait only for LDS traffic).  Same arithmetic, in
// the same order, as k_xcorr_lds.
template <int G, int PF, unsigned RM>
__global__ void __launch_bounds__(256)
k_xcorr_lds_pf(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
               FftDesc fdg, int n, int nhz, int nb, int C, int T, int ntiles, unsigned grp_mask, int lo, int w) {
    extern __shared__ float2 lds[];
    int b, tile;
    if (!xcd_tile(blockIdx.x, nb, ntiles, &tile, &b)) return;
    float2* A = lds;
    float2* Bf = lds + n * T;
    float2* acc = lds + 2 * n * T;                 // [G][n * T]
    const FftDesc fd = stage_twiddles(fdg, lds + (2 + G) * n * T);
    const int plane = n * nhz;
    const int j0 = tile * T;
    const int nt = min(T, plane - j0);
    const int ne = n * nt;
    const float ntinv = 1.0f / (float)nt;
    // per-thread element offsets in a channel plane (-1: none); bit 30 marks
    // slabs outside the support box (zero, not loaded)
    int off[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int e = threadIdx.x + u * 256;
        off[u] = -1;
        if (e < ne) {
            const int xx = fdiv(e, nt, ntinv), t = e - xx * nt;
            off[u] = xx * plane + t;
            if (xx < lo || xx >= lo + w) off[u] |= 1 << 30;
        }
    }
    for (int e = threadIdx.x; e < G * ne; e += 256) acc[e] = make_float2(0.f, 0.f);
    const float2* sb = spec + (size_t)b * C * n * (size_t)plane + j0;
    float2 sv[PF], tv[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        sv[u] = make_float2(0.f, 0.f);
        if (off[u] >= 0 && !(off[u] & (1 << 30))) sv[u] = sb[off[u]];
    }
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int u = 0; u < PF; ++u)
            if (off[u] >= 0) A[threadIdx.x + u * 256] = sv[u];
        const float2* tc = tgt + (size_t)c * n * plane + j0;
#pragma unroll
        for (int u = 0; u < PF; ++u)
            if (off[u] >= 0) tv[u] = tc[off[u] & ~(1 << 30)];
        if (c + 1 < C) {
            const float2* sn = sb + (size_t)(c + 1) * n * plane;
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                sv[u] = make_float2(0.f, 0.f);
                if (off[u] >= 0 && !(off[u] & (1 << 30))) sv[u] = sn[off[u]];
            }
        }
        __syncthreads();
        const float2* r = fft_lds_pp<RM>(A, Bf, fd, nt, nt, 1, true);
        float2* ag = acc + ((grp_mask >> c) & 1) * ne;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * 256;
            if (off[u] >= 0) ag[e] = cadd(ag[e], cmul(r[e], tv[u]));
        }
        __syncthreads();
    }
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * 256;
            if (off[u] >= 0) A[e] = cconj(acc[g * ne + e]);
        }
        __syncthreads();
        const float2* r = fft_lds_pp<RM>(A, Bf, fd, nt, nt, 1, true);
        float2* dst = out + ((size_t)b * G + g) * n * (size_t)plane + j0;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * 256;
            if (off[u] >= 0) dst[off[u] & ~(1 << 30)] = cconj(r[e]);
        }
        __syncthreads();
    }
}

// Register-accumulating form of k_xcorr_lds_pf: each thread owns the same
// tile elements in every phase, so the group sum lives in registers and the
// workgroup needs two tile buffers of LDS instead of 2 + G (more workgroups
// per CU).  The groups run one after the other (a group's channels in
// channel order, then its inverse), so one accumulator set serves every
// group: the same arithmetic in the same order as accumulating all groups
// at once.
// 5 waves per SIMD (96 VGPRs) hold without spills when the tile values per
// thread are at most 6 and the radices are small: radix 11 or 13 stages, or
// radix 7 and 8 together, need more registers and keep the compiler's choice.
// The one-group 4-column compile-time tiles of 343 = 7 7 7, 350 = 2 5 5 7 and
// 378 = 2 3 3 3 7 spill 12-36 VGPRs at that budget and keep the compiler's
// choice too.
constexpr int xcorr_waves_hint(int PF, unsigned mask, int G = 0, int n = 0, int nl = 0) {
    return (PF <= 6 && (mask & (radix_bit(11) | radix_bit(13))) == 0 &&
            (mask & (radix_bit(7) | radix_bit(8))) != (radix_bit(7) | radix_bit(8)) &&
            !(G == 1 && nl == 4 && (n == 343 || n == 350 || n == 378)))
               ? 5
               : 1;
}

template <int G, int PF, unsigned RM, int NT, class SP = DynamicTile>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(xcorr_waves_hint(PF, SP::kStatic ? SP::kMask : RM, G, SP::kN, SP::kNL))))
k_xcorr_reg_pf(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
               FftDesc fdg, int n_, int nhz_, int nb, int C, int T_, int ntiles, unsigned grp_mask, int lo, int w,
               int tile0) {
    // a compile-time tile plan fixes the grid edge and the tile width
    const int n = SP::kStatic ? SP::kN : n_;
    const int nhz = SP::kStatic ? SP::kN / 2 + 1 : nhz_;
    const int T = SP::kStatic ? SP::kNL : T_;
    // static plans: every tile is full, so slots below kFull always hold an element
    constexpr int kFull = SP::kStatic ? SP::kN * SP::kNL : 0;
    extern __shared__ float2 lds[];
    int b, tile;
    if (!xcd_tile(blockIdx.x, nb, ntiles, &tile, &b)) return;
    tile += tile0;   // tiles [tile0, tile0 + ntiles) of the plane
    float2* A = lds;
    float2* Bf = lds + n * T;
    const FftDesc fd = stage_twiddles(fdg, lds + 2 * n * T);
    const int plane = n * nhz;
    const int j0 = tile * T;
    const int nt = SP::kStatic ? T : min(T, plane - j0);
    const int ne = n * nt;
    const float ntinv = 1.0f / (float)nt;
    // Byte offset of each slot's element (x, j0 + t) from row lo of its
    // column tile: ((x - lo) plane + t) * 8, wrapping below row lo; slots
    // without an element hold kNoSlot.  Loads go through buffer resources
    // without branches: the query's resource spans the support box's rows,
    // so rows outside it read as zero; the target's the whole channel (its
    // offsets are these plus lo * plane * 8, which kNoSlot keeps out of range).
    constexpr uint32_t kNoSlot = 0x80000000u;
    uint32_t boff[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int e = threadIdx.x + u * NT;
        boff[u] = kNoSlot;
        if ((u + 1) * NT <= kFull || e < ne) {
            const int xx = qdiv<SP::kStatic>(e, nt, ntinv), t = e - xx * nt;
            boff[u] = (uint32_t)(((xx - lo) * plane + t) * (int)sizeof(float2));
        }
    }
    const uint32_t tshift = (uint32_t)(lo * plane * (int)sizeof(float2));
#define VX_SLOT(u) ((u + 1) * NT <= kFull || boff[u] != kNoSlot)
    // channel order: group 0's channels, then group 1's (G == 2)
    auto in_group = [&](int c, int g) { return G == 1 || (int)((grp_mask >> c) & 1) == g; };
    auto next_channel = [&](int c, int g) {   // next channel of group g after c, else -1
        for (int k = c + 1; k < C; ++k)
            if (in_group(k, g)) return k;
        return -1;
    };
    const float2* sb = spec + (size_t)b * C * n * (size_t)plane + j0;
    const uint32_t box_bytes = (uint32_t)(w * plane * (int)sizeof(float2));
    const uint32_t chan_bytes = (uint32_t)((n * plane - j0) * (int)sizeof(float2));
    auto load_query = [&](int ch, float2 (&v)[PF]) {
        const auto rs = buf_rsrc(sb + ((size_t)ch * n + lo) * plane, box_bytes);
#pragma unroll
        for (int u = 0; u < PF; ++u) v[u] = buf_load_f2(rs, boff[u]);
    };
    float2 sv[PF], tv[PF], acc[PF];
    int c = next_channel(-1, 0);
    int gc = 0;   // group of the prefetched channel
    if (c < 0 && G == 2) { c = next_channel(-1, 1); gc = 1; }
    if (c >= 0) load_query(c, sv);
#pragma unroll 1
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int u = 0; u < PF; ++u) acc[u] = make_float2(0.f, 0.f);
        for (; c >= 0 && gc == g;) {
            // the channel after this one: the next of this group, else the
            // first of the next group
            int pc = next_channel(c, g), pg = g;
            if (pc < 0 && g + 1 < G) { pc = next_channel(-1, g + 1); pg = g + 1; }
#pragma unroll
            for (int u = 0; u < PF; ++u)
                if (VX_SLOT(u)) A[threadIdx.x + u * NT] = sv[u];
            {
                const auto rt = buf_rsrc(tgt + (size_t)c * n * plane + j0, chan_bytes);
#pragma unroll
                for (int u = 0; u < PF; ++u) tv[u] = buf_load_f2(rt, boff[u] + tshift);
            }
            if (pc >= 0) load_query(pc, sv);
            __syncthreads();
            const float2* r = tile_fft<SP, RM, NT>(A, Bf, fd, nt);
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int e = threadIdx.x + u * NT;
                if (VX_SLOT(u)) acc[u] = cadd(acc[u], cmul(r[e], tv[u]));
            }
            __syncthreads();
            c = pc;
            gc = pg;
        }
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * NT;
            if (VX_SLOT(u)) A[e] = cconj(acc[u]);
        }
        __syncthreads();
        const float2* r = tile_fft<SP, RM, NT>(A, Bf, fd, nt);
        float2* dst = out + ((size_t)b * G + g) * n * (size_t)plane + j0;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int e = threadIdx.x + u * NT;
            if (VX_SLOT(u)) dst[(boff[u] + tshift) / sizeof(float2)] = cconj(r[e]);
        }
        __syncthreads();
    }
#undef VX_SLOT
}

// ---------------------------------------------------------------------------
// x pass on MFMA: the x-axis DFTs as real-form complex matrix products with
// split bf16 operands (v = hi + lo, hi = bf16(v), lo = bf16(v - hi)); each
// product is hi*hi + hi*lo + lo*hi accumulated in fp32 (the dropped lo*lo
// term is ~2^-16 of the product).  Per (tile of 32 or 64 columns, rotation):
//   Y_c = F X_c          F [2NP x 2KP]: rows (Re, Im) of kx, columns
//                         (Re, Im) of the query slabs x in [lo, lo + KP)
//   P   = sum_c Y_c * tgt_c             (fp32 VALU; tgt holds conj spectra)
//   Z   = G P            G [2NP x 2NP]: unnormalised inverse DFT
// Wave v of NW owns the row tiles i = v, v + NW, ... (Re rows, and their Im
// rows i + NPT) for all columns, so no wave repeats another's A operand.  The
// matrices are stored pre-swizzled in MFMA fragment order (one 1 KB run per
// 16x32 tile, 16 bytes per lane, read straight from L2); the data operand is
// staged in LDS as [column][k] split bf16 with a 16-byte row pad.
// v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][k], B[k][l % 16] for its
// eight k (the same eight in both operands, so any fixed assignment of k to
// (l / 16, v) gives the product; here k = 8 (l / 16) + v), and
// D[4 (l / 16) + v][l % 16], v < 4.
// ---------------------------------------------------------------------------

typedef float floatx4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void split_bf16(float v, unsigned short& hi, unsigned short& lo) {
    hi = bf16_rne(v);
    lo = bf16_rne(v - __uint_as_float((uint32_t)hi << 16));
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMfmaPad = 8;        // LDS row padding (bf16 elements, 16 bytes)
constexpr int kMfmaMaxNPT = 12;    // row tiles of 16: grid edges up to 192

__device__ __forceinline__ bf16x8_t ld_bf16x8(const unsigned short* p) {
    return __builtin_bit_cast(bf16x8_t, *(const u32x4_t*)p);
}

// acc[p][h][j] += A(row tile i + h NPT, k step ks) * B(column tile j) with
// split operands, on v_mfma_f32_16x16x32_bf16 (k = 8 (lane / 16) + v in both
// operands).  A: fragment-ordered matrix, one 1 KB run per (row tile, k step);
// B: LDS [column][k] rows of `stride` elements.
template <int MAXP, int NPT, int CT, int NW>
__device__ __forceinline__ void mfma_split_step(floatx4_t (&acc)[MAXP][2][CT], const unsigned short* __restrict__ Ah,
                                                const unsigned short* __restrict__ Al, int KS, int ks,
                                                const unsigned short* bhp, const unsigned short* blp, int stride,
                                                int wave, int lane) {
    const int l16 = lane & 15, l8 = (lane >> 4) * 8;
    bf16x8_t bh[CT], bl[CT];
#pragma unroll
    for (int j = 0; j < CT; ++j) {
        bh[j] = ld_bf16x8(&bhp[(16 * j + l16) * stride + 32 * ks + l8]);
        bl[j] = ld_bf16x8(&blp[(16 * j + l16) * stride + 32 * ks + l8]);
    }
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
        const int i = wave + NW * p;
        if (i >= NPT) break;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const size_t ro = ((size_t)((i + h * NPT) * KS + ks) * 64 + lane) * 8;
            const bf16x8_t ah = ld_bf16x8(&Ah[ro]);
            const bf16x8_t al = ld_bf16x8(&Al[ro]);
#pragma unroll
            for (int j = 0; j < CT; ++j) {
                acc[p][h][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[p][h][j], 0, 0, 0);
                acc[p][h][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[p][h][j], 0, 0, 0);
                acc[p][h][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[p][h][j], 0, 0, 0);
            }
        }
    }
}

template <int NPT, int CT, int NW>
__global__ void __launch_bounds__(64 * NW)
k_xcorr_mfma(const float2* __restrict__ spec, const float2* __restrict__ tgt, float2* __restrict__ out,
             const unsigned short* __restrict__ Fh, const unsigned short* __restrict__ Fl,
             const unsigned short* __restrict__ Gh, const unsigned short* __restrict__ Gl,
             int n, int nhz, int KP, int lo, int w, int C, int nb, int ntiles) {
    constexpr int COLS = 16 * CT;
    constexpr int NP = 16 * NPT;
    constexpr int NP2 = 2 * NP;
    constexpr int MAXP = (NPT + NW - 1) / NW;
    constexpr int NT = 64 * NW;
    extern __shared__ unsigned short sm[];
    int b, tile;
    if (!xcd_tile(blockIdx.x, nb, ntiles, &tile, &b)) return;
    const int plane = n * nhz;
    const int j0 = tile * COLS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, l4 = (lane >> 4) * 4;
    const int K2 = 2 * KP, KS = K2 / 32;
    const int SK = K2 + kMfmaPad;
    unsigned short* xh = sm;
    unsigned short* xl = sm + COLS * SK;
    float P[MAXP][CT][4][2];
#pragma unroll
    for (int p = 0; p < MAXP; ++p)
#pragma unroll
        for (int j = 0; j < CT; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) { P[p][j][v][0] = 0.f; P[p][j][v][1] = 0.f; }
    for (int c = 0; c < C; ++c) {
        // stage X_c^T: column t, k < KP -> Re of slab lo + k, k >= KP -> Im
        const float2* src = spec + ((size_t)b * C + c) * n * (size_t)plane + j0;
        for (int e = threadIdx.x; e < COLS * KP; e += NT) {
            const int k = e / COLS, t = e - k * COLS;
            float2 v = make_float2(0.f, 0.f);
            if (k < w && j0 + t < plane) v = src[(size_t)(lo + k) * plane + t];
            unsigned short h, l;
            split_bf16(v.x, h, l);
            xh[t * SK + k] = h; xl[t * SK + k] = l;
            split_bf16(v.y, h, l);
            xh[t * SK + KP + k] = h; xl[t * SK + KP + k] = l;
        }
        __syncthreads();
        floatx4_t acc[MAXP][2][CT];
#pragma unroll
        for (int p = 0; p < MAXP; ++p)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int j = 0; j < CT; ++j) acc[p][h][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < KS; ++ks)
            mfma_split_step<MAXP, NPT, CT, NW>(acc, Fh, Fl, KS, ks, xh, xl, SK, wave, lane);
        // P += Y_c * tgt_c
        const float2* tc = tgt + (size_t)c * n * plane + j0;
#pragma unroll
        for (int p = 0; p < MAXP; ++p) {
            const int i = wave + NW * p;
            if (i >= NPT) break;
#pragma unroll
            for (int j = 0; j < CT; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int kx = 16 * i + l4 + v, t = 16 * j + l16;
                    if (kx < n && j0 + t < plane) {
                        const float2 tv = tc[(size_t)kx * plane + t];
                        const float yr = acc[p][0][j][v], yi = acc[p][1][j][v];
                        P[p][j][v][0] += yr * tv.x - yi * tv.y;
                        P[p][j][v][1] += yr * tv.y + yi * tv.x;
                    }
                }
        }
        __syncthreads();
    }
    // stage P^T as the inverse's data operand: column t, r < NP -> Re kx = r, else Im
    const int SP = NP2 + kMfmaPad;
    unsigned short* ph = sm;
    unsigned short* pl = sm + COLS * SP;
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
        const int i = wave + NW * p;
        if (i >= NPT) break;
#pragma unroll
        for (int j = 0; j < CT; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kx = 16 * i + l4 + v, t = 16 * j + l16;
                unsigned short h, l;
                split_bf16(P[p][j][v][0], h, l);
                ph[t * SP + kx] = h; pl[t * SP + kx] = l;
                split_bf16(P[p][j][v][1], h, l);
                ph[t * SP + NP + kx] = h; pl[t * SP + NP + kx] = l;
            }
    }
    __syncthreads();
    floatx4_t acc[MAXP][2][CT];
#pragma unroll
    for (int p = 0; p < MAXP; ++p)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < CT; ++j) acc[p][h][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < NP2 / 32; ++ks)
        mfma_split_step<MAXP, NPT, CT, NW>(acc, Gh, Gl, NP2 / 32, ks, ph, pl, SP, wave, lane);
    float2* dst = out + (size_t)b * n * (size_t)plane + j0;
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
        const int i = wave + NW * p;
        if (i >= NPT) break;
#pragma unroll
        for (int j = 0; j < CT; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int x = 16 * i + l4 + v, t = 16 * j + l16;
                if (x < n && j0 + t < plane)
                    dst[(size_t)x * plane + t] = make_float2(acc[p][0][j][v], acc[p][1][j][v]);
            }
    }
}

// Fragment-ordered split bf16 DFT matrices of the MFMA x pass (host side).
// rows 2 NP, columns K2 (= 2 KP forward, 2 NP inverse); element (r, k) at
// ((r / 16 * KS + k / 32) * 64 + lane) * 8 + k % 8, lane = r % 16 + 16 ((k % 32) / 8),
// KS = K2 / 32.  fwd: F (columns are query slabs x = lo + k); else G.
static void build_dft_matrix(int n, int NP, int KP, int lo, int w, bool fwd,
                             std::vector<unsigned short>& hi, std::vector<unsigned short>& lov) {
    const int R2 = 2 * NP, K2 = fwd ? 2 * KP : 2 * NP, KS = K2 / 32, half = K2 / 2;
    hi.assign((size_t)R2 * K2, 0);
    lov.assign((size_t)R2 * K2, 0);
    auto to_bf16 = [](float f) {
        uint32_t u;
        memcpy(&u, &f, 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    };
    for (int r = 0; r < R2; ++r) {
        const int kx = r % NP;
        const bool im_row = r >= NP;
        if (kx >= n) continue;
        for (int k = 0; k < K2; ++k) {
            const int m = k % half;
            const bool im_col = k >= half;
            const int x = fwd ? lo + m : m;
            if ((fwd && m >= w) || x >= n) continue;
            const double ang = 2.0 * M_PI * (double)(((long long)kx * x) % n) / n;
            const double cs = cos(ang), sn = sin(ang);
            // fwd: Y = sum X e^{-i ang}; inverse: Z = sum P e^{+i ang}
            double val;
            if (!im_row) val = im_col ? (fwd ? sn : -sn) : cs;
            else         val = im_col ? cs : (fwd ? -sn : sn);
            const float f = (float)val;
            const unsigned short h = to_bf16(f);
            uint32_t hu = (uint32_t)h << 16;
            float hf;
            memcpy(&hf, &hu, 4);
            const unsigned short l = to_bf16(f - hf);
            const int lane = (r % 16) + 16 * ((k % 32) / 8);
            const size_t idx = ((size_t)((r / 16) * KS + k / 32) * 64 + lane) * 8 + (k % 8);
            hi[idx] = h;
            lov[idx] = l;
        }
    }
}

// x-pass MFMA shapes: tile = 32 columns (4 waves) or 64 columns (8 waves)
static void mfma_shape(int tile, int* ct, int* nw) {
    if (tile == 32) { *ct = 2; *nw = 4; }
    else            { *ct = 4; *nw = 8; }
}

static size_t xcorr_mfma_lds(int NP, int KP, int tile) {
    int ct, nw;
    mfma_shape(tile, &ct, &nw);
    const int s = std::max(2 * KP, 2 * NP) + kMfmaPad;
    return (size_t)16 * ct * s * 2 * sizeof(unsigned short);
}

template <int NPT, int CT, int NW>
static void launch_xcorr_mfma_t(dim3 grid, size_t lds, hipStream_t st, const float2* spec, const float2* tgt,
                                float2* out, const unsigned short* const* mats, int n, int nhz, int KP,
                                int lo, int w, int C, int nb, int ntiles) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_xcorr_mfma<NPT, CT, NW>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    hipLaunchKernelGGL((k_xcorr_mfma<NPT, CT, NW>), grid, dim3(64 * NW), lds, st, spec, tgt, out, mats[0], mats[1],
                       mats[2], mats[3], n, nhz, KP, lo, w, C, nb, ntiles);
}

int launch_xcorr_mfma(const float2* spec, const float2* tgt, float2* out, const unsigned short* const* mats,
                      int n, int nb, int C, int KP, int lo, int w, int tile, hipStream_t st) {
    const int nhz = n / 2 + 1;
    const int NPT = (n + 15) / 16;
    VX_REQUIRE(NPT <= kMfmaMaxNPT, VX_ERR_UNSUPPORTED, "grid edge %d too large for the MFMA x pass", n);
    VX_REQUIRE(KP % 16 == 0 && KP >= w && KP <= 16 * NPT, VX_ERR_STATE, "bad MFMA x-pass width");
    int ct, nw;
    mfma_shape(tile, &ct, &nw);
    const size_t lds = xcorr_mfma_lds(16 * NPT, KP, tile);
    VX_REQUIRE(lds <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d too large for the MFMA x pass", n);
    const int cols = 16 * ct;
    const int ntiles = (n * nhz + cols - 1) / cols;
    dim3 grid((unsigned)(((ntiles + 7) / 8) * 8 * nb));
    if (ct == 2) {
        switch (NPT) {
#define VX_NPT(k) case k: launch_xcorr_mfma_t<k, 2, 4>(grid, lds, st, spec, tgt, out, mats, n, nhz, KP, lo, w, C, nb, ntiles); break;
            VX_NPT(1) VX_NPT(2) VX_NPT(3) VX_NPT(4) VX_NPT(5) VX_NPT(6)
            VX_NPT(7) VX_NPT(8) VX_NPT(9) VX_NPT(10) VX_NPT(11) VX_NPT(12)
#undef VX_NPT
            default: break;
        }
    } else {
        switch (NPT) {
#define VX_NPT(k) case k: launch_xcorr_mfma_t<k, 4, 8>(grid, lds, st, spec, tgt, out, mats, n, nhz, KP, lo, w, C, nb, ntiles); break;
            VX_NPT(1) VX_NPT(2) VX_NPT(3) VX_NPT(4) VX_NPT(5) VX_NPT(6)
            VX_NPT(7) VX_NPT(8) VX_NPT(9) VX_NPT(10) VX_NPT(11) VX_NPT(12)
#undef VX_NPT
            default: break;
        }
    }
    VX_LAUNCH_CHECK();
    return VX_OK;
}

// Inverse: sum[bg][x] ([ky][kz]) -> real slab ([y][z]) -> argmax key.
// Optionally writes the real correlation (prob-mode mixing).
template <int MAXV, int NT, unsigned RM = 0xffffffffu, class SP = DynamicPlan>
__global__ void __launch_bounds__(NT)
k_inv_yz(const float2* __restrict__ sum, FftDesc fdg, int n_, int nhz_, int G,
         unsigned long long* __restrict__ keys, int nkeys, float* __restrict__ real_out) {
    const int n = SP::kStatic ? SP::kN : n_;
    const int nhz = SP::kStatic ? SP::kN / 2 + 1 : nhz_;
    extern __shared__ float2 lds[];
    const FftDesc fd = stage_twiddles(fdg, lds + n * nhz);
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    const int x = blockIdx.x;
    const int bg = blockIdx.y;
    const int nthr = blockDim.x;
    const float2* src = sum + ((size_t)bg * n + x) * (size_t)n * nhz;
    {
        float2 v[MAXV];
#pragma unroll
        for (int m = 0; m < MAXV; ++m) {
            const int e = threadIdx.x + m * nthr;
            if (e < n * nhz) v[m] = src[e];
        }
#pragma unroll
        for (int m = 0; m < MAXV; ++m) {
            const int e = threadIdx.x + m * nthr;
            if (e < n * nhz) lds[e] = cconj(v[m]);
        }
    }
    __syncthreads();
    slab_fft_tile<SP, MAXV, RM, NT>(lds, fd, nhz);   // conj of the y-inverse
    const int npair = (n + 1) >> 1;
    const int half = n >> 1;
    const bool even = (n & 1) == 0;
    const float ninv_p = 1.0f / (float)n;
    // pack conj(Z), Z[k] = A[k] + i B[k] with A, B = rows 2l, 2l+1 after the
    // y-inverse (lds holds r = conj of them); Hermitian extension for k > n/2;
    // imaginary parts of the DC (and Nyquist) bins dropped as c2r does.
    {
        constexpr int MU = MAXV;
        float2 Zc[MU];
        int dst[MU];
        const int nu = npair * n;
#pragma unroll
        for (int m = 0; m < MU; ++m) {
            const int u = threadIdx.x + m * nthr;
            dst[m] = -1;
            if (u < nu) {
                const int l = qdiv<SP::kStatic>(u, n, ninv_p), k = u - l * n;
                const bool hasb = (2 * l + 1) < n;
                if (k <= half) {
                    float2 ra = lds[(2 * l) * nhz + k];
                    float2 rb = hasb ? lds[(2 * l + 1) * nhz + k] : make_float2(0.f, 0.f);
                    if (k == 0 || (even && k == half)) { ra.y = 0.f; rb.y = 0.f; }
                    Zc[m] = make_float2(ra.x + rb.y, ra.y - rb.x);
                } else {
                    const int kk = n - k;
                    float2 ra = lds[(2 * l) * nhz + kk];
                    float2 rb = hasb ? lds[(2 * l + 1) * nhz + kk] : make_float2(0.f, 0.f);
                    Zc[m] = make_float2(ra.x - rb.y, -ra.y - rb.x);
                }
                dst[m] = l * n + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < MU; ++m)
            if (dst[m] >= 0) lds[dst[m]] = Zc[m];
        __syncthreads();
    }
    slab_fft_lines<SP, MAXV, RM, NT, SP::kStatic ? (SP::kN + 1) / 2 : 0>(lds, fd, npair);
    // z[n] = conj(F): row 2l = F.x, row 2l+1 = -F.y
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    const size_t n3 = (size_t)n * n * n;
    for (int e = threadIdx.x; e < npair * n; e += nthr) {
        const int l = qdiv<SP::kStatic>(e, n, ninv_p), z = e - l * n;
        const float2 f = lds[e];
        const uint32_t ia = ((uint32_t)x * n + 2 * l) * n + z;
        const float va = f.x;
        if (better(va, ia, bv, bi)) { bv = va; bi = ia; }
        if (real_out) real_out[(size_t)bg * n3 + ia] = va;
        if (2 * l + 1 < n) {
            const uint32_t ib = ia + n;
            const float vb = -f.y;
            if (better(vb, ib, bv, bi)) { bv = vb; bi = ib; }
            if (real_out) real_out[(size_t)bg * n3 + ib] = vb;
        }
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) {
        const int b = bg / G, g = bg - b * G;
        atomicMax(&keys[(size_t)b * nkeys + g], argmax_key(bv, bi));
    }
}

// ---------------------------------------------------------------------------
// Pencil passes for grids whose (y, z) slab does not fit in LDS
// ---------------------------------------------------------------------------

// Forward r2c along z for a chunk of P line pairs of one x-slab: two real
// rows packed into one complex line, FFT, unpacked into half spectra.
// Only slabs x in [lo, lo + w), pairs covering rows [lo, lo + w) and loads of
// z in [lo, lo + w).
// With a compile-time plan (SP = StaticPlan<radices>) the lines transform in
// place on registers (at most kPencilMaxV values per thread).

template <class SP = DynamicPlan>
__global__ void __launch_bounds__(256)
k_fwd_z(const float* __restrict__ q, float2* __restrict__ spec, FftDesc fdg, int n_, int nhz_, int lo, int w, int P) {
    extern __shared__ float2 lds[];
    const int n = SP::kStatic ? SP::kN : n_;
    const int nhz = SP::kStatic ? SP::kN / 2 + 1 : nhz_;
    float2* A = lds;
    float2* Bf = lds + P * n;
    // the in-place compile-time FFT needs no ping-pong buffer: twiddles follow A
    const FftDesc fd = stage_twiddles(fdg, lds + (SP::kStatic ? 1 : 2) * P * n);
    const int x = lo + blockIdx.y;
    const size_t plane = blockIdx.z;
    const int l0 = (lo >> 1) + blockIdx.x * P;
    const int l1 = min(l0 + P, ((lo + w - 1) >> 1) + 1);
    const int np = l1 - l0;
    if (np <= 0) return;
    const float* src = q + (plane * n + x) * (size_t)n * n;
    const float ninv = 1.0f / (float)n;
    for (int e = threadIdx.x; e < np * n; e += 256) {
        const int ll = qdiv<SP::kStatic>(e, n, ninv), z = e - ll * n, l = l0 + ll;
        // the query planes are zero outside the support box: no loads there
        const bool in = z >= lo && z < lo + w;
        const float re = in ? src[(size_t)(2 * l) * n + z] : 0.f;
        const float im = (in && 2 * l + 1 < n) ? src[(size_t)(2 * l + 1) * n + z] : 0.f;
        A[e] = make_float2(re, im);
    }
    __syncthreads();
    const float2* r = A;
    if constexpr (SP::kStatic) SP::template lines_inplace<256, kPencilMaxV>(A, fd.twid, np);
    else r = fft_lds_pp(A, Bf, fd, np, 1, n, false);
    float2* out = spec + (plane * n + x) * (size_t)n * nhz;
    const float hinv = 1.0f / (float)nhz;
    for (int u = threadIdx.x; u < np * nhz; u += 256) {
        const int ll = qdiv<SP::kStatic>(u, nhz, hinv), k = u - ll * nhz, l = l0 + ll;
        const float2 zk = r[ll * n + k];
        const float2 wv = r[ll * n + (k == 0 ? 0 : n - k)];
        out[(size_t)(2 * l) * nhz + k] = make_float2(0.5f * (zk.x + wv.x), 0.5f * (zk.y - wv.y));
        if (2 * l + 1 < n) out[(size_t)(2 * l + 1) * nhz + k] = make_float2(0.5f * (zk.y + wv.y), 0.5f * (wv.x - zk.x));
    }
}

// In-place complex FFT along y of a tile of T kz-columns of one x-slab.
// Forward: rows outside [ylo, ylo + yw) are read as zero.  Inverse
// (unnormalised): conj on load and store.
// One workgroup per (column tile, x-slab, plane) over tiles [tile0, tile0 +
// ntile) and nxs x-slabs, on a 1-D grid mapped XCD-aware: workgroup ids reach
// the 8 XCDs round-robin, and XCD k takes a contiguous range of (tile, slab,
// plane), so the column tiles that share a row's cache lines (T columns are
// 8 T bytes of a row) read them through one L2.  With a compile-time tile
// plan (SP = StaticTile<T, radices>) every tile of the launch is full.
template <class SP = DynamicTile>
__global__ void __launch_bounds__(256)
k_axis_y(float2* __restrict__ buf, FftDesc fdg, int n_, int nhz_, int T_, int inverse, int ylo, int yw, int xlo,
         int ntile, int nxs, int total, int tile0) {
    extern __shared__ float2 lds[];
    const int n = SP::kStatic ? SP::kN : n_;
    const int nhz = SP::kStatic ? SP::kN / 2 + 1 : nhz_;
    const int T = SP::kStatic ? SP::kNL : T_;
    const int per = (total + 7) >> 3;
    const int m = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if ((int)(blockIdx.x >> 3) >= per || m >= total) return;
    const int tile = m % ntile + tile0, rest = m / ntile;
    const int xs = rest % nxs, plane = rest / nxs;
    float2* A = lds;
    float2* Bf = lds + n * T;
    const FftDesc fd = stage_twiddles(fdg, lds + 2 * n * T);
    const int k0 = tile * T;
    const int nt = SP::kStatic ? T : min(T, nhz - k0);
    const int x = xlo + xs;
    float2* base = buf + ((size_t)plane * n + x) * (size_t)n * nhz + k0;
    const float ntinv = 1.0f / (float)nt;
    for (int e = threadIdx.x; e < n * nt; e += 256) {
        const int y = qdiv<SP::kStatic>(e, nt, ntinv), t = e - y * nt;
        float2 v = make_float2(0.f, 0.f);
        if (y >= ylo && y < ylo + yw) v = base[(size_t)y * nhz + t];
        A[e] = inverse ? cconj(v) : v;
    }
    __syncthreads();
    const float2* r;
    if constexpr (SP::kStatic) r = SP::template tile_fft<256>(A, Bf, fd.twid);
    else r = fft_lds_pp(A, Bf, fd, nt, nt, 1, true);
    for (int e = threadIdx.x; e < n * nt; e += 256) {
        const int y = qdiv<SP::kStatic>(e, nt, ntinv), t = e - y * nt;
        base[(size_t)y * nhz + t] = inverse ? cconj(r[e]) : r[e];
    }
}

// Inverse c2r along z for a chunk of P line pairs (rows 2l, 2l+1 of the
// y-inverted half spectrum), then max + first argmax -> atomicMax.
template <class SP = DynamicPlan>
__global__ void __launch_bounds__(256)
k_inv_z(const float2* __restrict__ sum, FftDesc fdg, int n_, int nhz_, int G, int P,
        unsigned long long* __restrict__ keys, int nkeys, float* __restrict__ real_out) {
    extern __shared__ float2 lds[];
    const int n = SP::kStatic ? SP::kN : n_;
    const int nhz = SP::kStatic ? SP::kN / 2 + 1 : nhz_;
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    float2* A = lds;
    float2* Bf = lds + P * n;
    // the in-place compile-time FFT needs no ping-pong buffer: twiddles follow A
    const FftDesc fd = stage_twiddles(fdg, lds + (SP::kStatic ? 1 : 2) * P * n);
    const int x = blockIdx.y;
    const int bg = blockIdx.z;
    const int npair = (n + 1) >> 1;
    const int l0 = blockIdx.x * P;
    const int np = min(P, npair - l0);
    const float2* src = sum + ((size_t)bg * n + x) * (size_t)n * nhz;
    const int half = n >> 1;
    const bool even = (n & 1) == 0;
    const float ninv = 1.0f / (float)n;
    for (int u = threadIdx.x; u < np * n; u += 256) {
        const int ll = qdiv<SP::kStatic>(u, n, ninv), k = u - ll * n, l = l0 + ll;
        const bool hasb = (2 * l + 1) < n;
        const int kk = (k <= half) ? k : n - k;
        float2 a = src[(size_t)(2 * l) * nhz + kk];
        float2 bb = hasb ? src[(size_t)(2 * l + 1) * nhz + kk] : make_float2(0.f, 0.f);
        float2 zc;
        if (k <= half) {
            if (k == 0 || (even && k == half)) { a.y = 0.f; bb.y = 0.f; }
            zc = make_float2(a.x - bb.y, -a.y - bb.x);       // conj(A + iB)
        } else {
            zc = make_float2(a.x + bb.y, a.y - bb.x);        // conj(conj A' + i conj B')
        }
        A[u] = zc;
    }
    __syncthreads();
    const float2* r = A;
    if constexpr (SP::kStatic) SP::template lines_inplace<256, kPencilMaxV>(A, fd.twid, np);
    else r = fft_lds_pp(A, Bf, fd, np, 1, n, false);
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    const size_t n3 = (size_t)n * n * n;
    for (int e = threadIdx.x; e < np * n; e += 256) {
        const int ll = qdiv<SP::kStatic>(e, n, ninv), z = e - ll * n, l = l0 + ll;
        const float2 f = r[e];
        const uint32_t ia = ((uint32_t)x * n + 2 * l) * n + z;
        if (better(f.x, ia, bv, bi)) { bv = f.x; bi = ia; }
        if (real_out) real_out[(size_t)bg * n3 + ia] = f.x;
        if (2 * l + 1 < n) {
            const uint32_t ib = ia + n;
            if (better(-f.y, ib, bv, bi)) { bv = -f.y; bi = ib; }
            if (real_out) real_out[(size_t)bg * n3 + ib] = -f.y;
        }
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) {
        const int b = bg / G, g = bg - b * G;
        atomicMax(&keys[(size_t)b * nkeys + g], argmax_key(bv, bi));
    }
}

template <typename K>
static int allow_lds_any(K kernel, size_t bytes) {
    if (bytes > 65536) VX_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return VX_OK;
}

// The y pass over nxs x-slabs (from xlo) of nplanes planes: full column tiles
// on a compile-time tile plan where the grid edge has one (8-column tiles for
// 201..256, 4-column above), a last partial tile and the other edges on the
// runtime plan.
static int launch_axis_y(float2* buf, const FftDesc& fd, int n, int inverse, int ylo, int yw, int xlo, int nxs,
                         int nplanes, bool use_static, hipStream_t st, bool fused = true) {
    const int nhz = n / 2 + 1;
    const int T = pencil_T(n);
    const size_t ly = (2 * (size_t)n * T + fd.ntw) * sizeof(float2);
    VX_REQUIRE(ly <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d too large", n);
    const int ntile = (nhz + T - 1) / T, full = nhz / T;
    auto grid = [](int total) { return dim3((unsigned)(((total + 7) / 8) * 8)); };
    auto go = [&](auto sp, auto width) {
        using TL = typename decltype(sp)::template Tile<decltype(width)::value>;
        const int total = full * nxs * nplanes;
        allow_lds_any(k_axis_y<TL>, ly);
        hipLaunchKernelGGL(k_axis_y<TL>, grid(total), dim3(256), ly, st, buf, fd, n, nhz, T, inverse, ylo, yw, xlo,
                           full, nxs, total, 0);
    };
    bool hit = false;
    if (use_static && fused && full > 0) {
        int rc, nfull;
        hit = launch_axis_y_fused(buf, fd, n, inverse, ylo, yw, xlo, nxs, nplanes, st, &nfull, &rc);
        if (rc != VX_OK) return rc;
    }
    if (hit) {
    } else if (use_static && full > 0 && T == 8)
        hit = with_static_plan(fd, TileLengths{}, [&](auto sp) { go(sp, std::integral_constant<int, 8>{}); });
    else if (use_static && full > 0 && T == 4)
        hit = with_static_plan(fd, LargeLengths{}, [&](auto sp) { go(sp, std::integral_constant<int, 4>{}); });
    VX_LAUNCH_CHECK();
    const int t0 = hit ? full : 0;
    if (t0 < ntile) {
        const int nt = ntile - t0, total = nt * nxs * nplanes;
        allow_lds_any(k_axis_y<DynamicTile>, ly);
        hipLaunchKernelGGL(k_axis_y<DynamicTile>, grid(total), dim3(256), ly, st, buf, fd, n, nhz, T, inverse, ylo,
                           yw, xlo, nt, nxs, total, t0);
        VX_LAUNCH_CHECK();
    }
    return VX_OK;
}

int launch_fwd_pencil(const float* q, float2* spec, const FftDesc& fd, int n, int nplanes, int lo, int w,
                      hipStream_t st, bool use_static = true, bool fused = true) {
    const int nhz = n / 2 + 1;
    const int P = pencil_pairs(n);
    const size_t lz = (2 * (size_t)P * n + fd.ntw) * sizeof(float2);
    VX_REQUIRE(lz <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d too large", n);
    const int l0 = lo >> 1, l1 = ((lo + w - 1) >> 1) + 1;
    if (use_static && fused) {
        int rc;
        if (launch_fwd_z_fused(q, spec, fd, n, nplanes, lo, w, st, &rc)) {
            if (rc != VX_OK) return rc;
            return launch_axis_y(spec, fd, n, 0, 2 * l0, 2 * (l1 - l0), lo, w, nplanes, use_static, st, fused);
        }
    }
    const dim3 gz((l1 - l0 + P - 1) / P, w, nplanes);
    auto zgo = [&](auto sp) {
        using SP = decltype(sp);
        const size_t l = SP::kStatic ? ((size_t)P * n + fd.ntw) * sizeof(float2) : lz;
        allow_lds_any(k_fwd_z<SP>, l);
        hipLaunchKernelGGL(k_fwd_z<SP>, gz, dim3(256), l, st, q, spec, fd, n, nhz, lo, w, P);
    };
    if (!(use_static && n * P <= 256 * kPencilMaxV &&
          (with_static_plan(fd, TileLengths{}, zgo) || with_static_plan(fd, LargeLengths{}, zgo))))
        zgo(DynamicPlan{});
    VX_LAUNCH_CHECK();
    return launch_axis_y(spec, fd, n, 0, 2 * l0, 2 * (l1 - l0), lo, w, nplanes, use_static, st, fused);
}

int launch_inv_pencil(float2* sum, const FftDesc& fd, int n, int nbg, int G, unsigned long long* keys, int nkeys,
                      float* real_out, hipStream_t st, bool use_static = true, bool fused = true, int chunks = 0) {
    const int nhz = n / 2 + 1;
    int rc = launch_axis_y(sum, fd, n, 1, 0, n, 0, n, nbg, use_static, st, fused);
    if (rc != VX_OK) return rc;
    if (use_static && fused && launch_inv_z_fused(sum, fd, n, nbg, G, keys, nkeys, real_out, st, &rc, chunks))
        return rc;
    const int P = pencil_pairs(n);
    const size_t lz = (2 * (size_t)P * n + fd.ntw) * sizeof(float2);
    VX_REQUIRE(lz <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d too large", n);
    const int npair = (n + 1) / 2;
    const dim3 gz((npair + P - 1) / P, n, nbg);
    auto zgo = [&](auto sp) {
        using SP = decltype(sp);
        const size_t l = SP::kStatic ? ((size_t)P * n + fd.ntw) * sizeof(float2) : lz;
        allow_lds_any(k_inv_z<SP>, l);
        hipLaunchKernelGGL(k_inv_z<SP>, gz, dim3(256), l, st, sum, fd, n, nhz, G, P, keys, nkeys, real_out);
    };
    if (!(use_static && n * P <= 256 * kPencilMaxV &&
          (with_static_plan(fd, TileLengths{}, zgo) || with_static_plan(fd, LargeLengths{}, zgo))))
        zgo(DynamicPlan{});
    VX_LAUNCH_CHECK();
    return VX_OK;
}

// Prob-mode mixed score (find_best_trans_mixed, VESPER_1_prob_new.py:1530-1545)
// in float32 with float32-rounded constants, no FMA:
//   mixed = fl(fl(a * fl(fl(v - vave) / vstd)) + fl(b * fl(fl(p - pave) / pstd)))
__global__ void __launch_bounds__(256)
k_mixed(const float* __restrict__ real, size_t n3, float vave, float vstd, float pave,
        float pstd, float wa, float wb, unsigned long long* __restrict__ keys, int nkeys) {
    __shared__ float s_v[16];
    __shared__ uint32_t s_i[16];
    const int b = blockIdx.y;
    const float* rv = real + (size_t)b * 2 * n3;
    const float* rp = rv + n3;
    float bv = -INFINITY;
    uint32_t bi = 0xffffffffu;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
         i += (size_t)gridDim.x * blockDim.x) {
        const float sv = __fdiv_rn(__fsub_rn(rv[i], vave), vstd);
        const float sp = __fdiv_rn(__fsub_rn(rp[i], pave), pstd);
        const float m = __fadd_rn(__fmul_rn(wa, sv), __fmul_rn(wb, sp));
        if (better(m, (uint32_t)i, bv, bi)) { bv = m; bi = (uint32_t)i; }
    }
    block_argmax(bv, bi, s_v, s_i);
    if (threadIdx.x == 0) atomicMax(&keys[(size_t)b * nkeys + 2], argmax_key(bv, bi));
}

// keys [nb][nkeys] -> best/trans rows [row0, row0 + nb), nout entries per row
// (nout <= nkeys: prob pass 1 reports {vec, prob} only).
__global__ void k_finalize(const unsigned long long* __restrict__ keys, int nb, int nkeys, int nout,
                           int n, int row0, float* __restrict__ best, int32_t* __restrict__ trans) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * nout) return;
    const int b = i / nout, g = i - b * nout;
    const unsigned long long k = keys[(size_t)b * nkeys + g];
    const uint32_t idx = 0xffffffffu - (uint32_t)(k & 0xffffffffu);
    const size_t o = (size_t)(row0 + b) * nout + g;
    best[o] = unordered_bits((uint32_t)(k >> 32));
    trans[o * 3 + 0] = (int32_t)(idx / ((uint32_t)n * n));
    trans[o * 3 + 1] = (int32_t)((idx / n) % n);
    trans[o * 3 + 2] = (int32_t)(idx % n);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

#define VX_SLAB_LAUNCH(KER, MV, NTH, ...)                                              \
    do {                                                                               \
        allow_lds(KER<MV, NTH>, lds);                                                  \
        hipLaunchKernelGGL((KER<MV, NTH>), grid, dim3(NTH), lds, st, __VA_ARGS__);     \
    } while (0)
#define VX_SLAB_LAUNCH_RM(KER, MV, NTH, RMV, ...)                                      \
    do {                                                                               \
        allow_lds(KER<MV, NTH, RMV>, lds);                                             \
        hipLaunchKernelGGL((KER<MV, NTH, RMV>), grid, dim3(NTH), lds, st, __VA_ARGS__); \
    } while (0)
#define VX_SLAB_LAUNCH_SP(KER, MV, NTH, RMV, SPT, ...)                                 \
    do {                                                                               \
        allow_lds(KER<MV, NTH, RMV, SPT>, lds);                                        \
        hipLaunchKernelGGL((KER<MV, NTH, RMV, SPT>), grid, dim3(NTH), lds, st, __VA_ARGS__); \
    } while (0)

int launch_fwd_zy(const float* q, float2* spec, const FftDesc& fd, int n, int nplanes, int lo, int w,
                  hipStream_t st, int nt_pref = 0, bool use_static = true, bool fused = true) {
    const int nhz = n / 2 + 1;
    int nthr, maxv;
    slab_shape(n, nhz, &nthr, &maxv, nt_pref);
    const size_t lds = ((size_t)n * nhz + fd.ntw) * sizeof(float2);
    VX_REQUIRE(maxv > 0 && lds <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d exceeds the LDS slab limit", n);
    if (use_static && fused) {
        int rc;
        if (launch_fwd_fused(q, spec, fd, n, nplanes, lo, w, st, nt_pref, &rc)) return rc;
    }
    dim3 grid(w, nplanes);
    if (maxv == 8 && use_static && with_static_plan(fd, SmallLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_SLAB_LAUNCH_SP(k_fwd_zy, 8, 512, SP::kMask, SP, q, spec, fd, n, nhz, lo, w);
        })) {
    } else if (maxv == 8) VX_SLAB_LAUNCH(k_fwd_zy, 8, 512, q, spec, fd, n, nhz, lo, w);
    else if (maxv == 17 && use_static && with_static_plan(fd, StaticLengths{}, [&](auto sp) {
                 using SP = decltype(sp);
                 VX_SLAB_LAUNCH_SP(k_fwd_zy, 16, 1024, SP::kMask, SP, q, spec, fd, n, nhz, lo, w);
             })) {
    } else if (maxv == 17 && fd.radix_mask == kRadix57)
        VX_SLAB_LAUNCH_RM(k_fwd_zy, 16, 1024, kRadix57, q, spec, fd, n, nhz, lo, w);
    else if (maxv == 17) VX_SLAB_LAUNCH(k_fwd_zy, 16, 1024, q, spec, fd, n, nhz, lo, w);
    else if (maxv == 16) VX_SLAB_LAUNCH(k_fwd_zy, 16, 512, q, spec, fd, n, nhz, lo, w);
    else if (maxv == 32) VX_SLAB_LAUNCH(k_fwd_zy, 32, 512, q, spec, fd, n, nhz, lo, w);
    else if (maxv == 24 && use_static && with_static_plan(fd, MediumLengths{}, [&](auto sp) {
                 using SP = decltype(sp);
                 VX_SLAB_LAUNCH_SP(k_fwd_zy, 24, 1024, SP::kMask, SP, q, spec, fd, n, nhz, lo, w);
             })) {
    } else               VX_SLAB_LAUNCH(k_fwd_zy, 24, 1024, q, spec, fd, n, nhz, lo, w);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

int launch_inv_yz(const float2* sum, const FftDesc& fd, int n, int nbg, int G,
                  unsigned long long* keys, int nkeys, float* real_out, hipStream_t st, int nt_pref = 0,
                  bool use_static = true, bool fused = true) {
    const int nhz = n / 2 + 1;
    int nthr, maxv;
    slab_shape(n, nhz, &nthr, &maxv, nt_pref);
    const size_t lds = ((size_t)n * nhz + fd.ntw) * sizeof(float2);
    VX_REQUIRE(maxv > 0 && lds <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d exceeds the LDS slab limit", n);
    if (use_static && fused) {
        int rc;
        if (launch_inv_fused(sum, fd, n, nbg, G, keys, nkeys, real_out, st, nt_pref, &rc)) return rc;
    }
    dim3 grid(n, nbg);
    if (maxv == 8 && use_static && with_static_plan(fd, SmallLengths{}, [&](auto sp) {
            using SP = decltype(sp);
            VX_SLAB_LAUNCH_SP(k_inv_yz, 8, 512, SP::kMask, SP, sum, fd, n, nhz, G, keys, nkeys, real_out);
        })) {
    } else if (maxv == 8) VX_SLAB_LAUNCH(k_inv_yz, 8, 512, sum, fd, n, nhz, G, keys, nkeys, real_out);
    else if (maxv == 17 && use_static && with_static_plan(fd, StaticLengths{}, [&](auto sp) {
                 using SP = decltype(sp);
                 VX_SLAB_LAUNCH_SP(k_inv_yz, 16, 1024, SP::kMask, SP, sum, fd, n, nhz, G, keys, nkeys, real_out);
             })) {
    } else if (maxv == 17 && fd.radix_mask == kRadix57)
        VX_SLAB_LAUNCH_RM(k_inv_yz, 16, 1024, kRadix57, sum, fd, n, nhz, G, keys, nkeys, real_out);
    else if (maxv == 17) VX_SLAB_LAUNCH(k_inv_yz, 16, 1024, sum, fd, n, nhz, G, keys, nkeys, real_out);
    else if (maxv == 16) VX_SLAB_LAUNCH(k_inv_yz, 16, 512, sum, fd, n, nhz, G, keys, nkeys, real_out);
    else if (maxv == 32) VX_SLAB_LAUNCH(k_inv_yz, 32, 512, sum, fd, n, nhz, G, keys, nkeys, real_out);
    else if (maxv == 24 && use_static && with_static_plan(fd, MediumLengths{}, [&](auto sp) {
                 using SP = decltype(sp);
                 VX_SLAB_LAUNCH_SP(k_inv_yz, 24, 1024, SP::kMask, SP, sum, fd, n, nhz, G, keys, nkeys, real_out);
             })) {
    } else               VX_SLAB_LAUNCH(k_inv_yz, 24, 1024, sum, fd, n, nhz, G, keys, nkeys, real_out);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

static int xcorr_T(int n) {
    // n * T <= 256 threads * 16 values
    int T = 4096 / n;
    if (T > 64) T = 64;
    if (T < 1) T = 1;
    return T;
}

constexpr int kXcorrMaxE = 12;   // tile values per thread in the x pass
constexpr int kXcorrPF = 8;      // prefetched tile values per thread (LDS-accumulating x pass)

static int xcorr_pp_T(int n) {
    // n * T <= 256 threads * kXcorrMaxE values; T <= 16 keeps two tiles of LDS small
    int T = 256 * kXcorrMaxE / n;
    if (T > 16) T = 16;
    if (T < 1) T = 1;
    return T;
}

// Radix set of a plan's length (bit r set when a radix-r stage occurs).

static unsigned radix_mask(const FftDesc& fd) {
    return fd.radix_mask;
}

template <int G, unsigned RM>
static void launch_pf(dim3 grid, size_t lds, hipStream_t st, const float2* spec, const float2* tgt, float2* out,
                      const FftDesc& fd, int n, int nhz, int nb, int C, int T, int ntiles, unsigned grp_mask, int lo,
                      int w) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_xcorr_lds_pf<G, kXcorrPF, RM>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_xcorr_lds_pf<G, kXcorrPF, RM>), grid, dim3(256), lds, st, spec, tgt, out, fd, n, nhz, nb,
                       C, T, ntiles, grp_mask, lo, w);
}

template <int G, unsigned RM, int NT, int PF = kXcorrPF, class SP = DynamicTile>
static void launch_reg_pf_nt(dim3 grid, size_t lds, hipStream_t st, const float2* spec, const float2* tgt,
                             float2* out, const FftDesc& fd, int n, int nhz, int nb, int C, int T, int ntiles,
                             unsigned grp_mask, int lo, int w, int tile0 = 0) {
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_xcorr_reg_pf<G, PF, RM, NT, SP>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_xcorr_reg_pf<G, PF, RM, NT, SP>), grid, dim3(NT), lds, st, spec, tgt, out, fd, n, nhz,
                       nb, C, T, ntiles, grp_mask, lo, w, tile0);
}

// grid of the x pass over ntiles tiles of nb rows (whole XCD groups of 8)
static dim3 xcorr_grid(int ntiles, int nb) { return dim3((unsigned)(((ntiles + 7) / 8) * 8 * nb)); }

// Runtime-plan form over tiles [tile0, tile0 + ntiles): the smallest
// workgroup (64, 128 or 256 threads) holding a tile of n * T values at
// kXcorrPF per thread, so tiles of few columns run as one- or two-wave
// workgroups whose barriers stall only their own waves; 256-thread tiles of
// at most 1536 values take 6 values per thread (fewer registers: 5 waves per
// SIMD instead of 4).
template <int G, unsigned RM>
static void launch_reg_pf_dyn(dim3 grid, size_t lds, hipStream_t st, const float2* spec, const float2* tgt,
                              float2* out, const FftDesc& fd, int n, int nhz, int nb, int C, int T, int ntiles,
                              unsigned grp_mask, int lo, int w, int tile0) {
    if (n * T <= 64 * kXcorrPF)
        launch_reg_pf_nt<G, RM, 64>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w,
                                    tile0);
    else if (n * T <= 128 * kXcorrPF)
        launch_reg_pf_nt<G, RM, 128>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w,
                                     tile0);
    else if (n * T <= 256 * 6)
        launch_reg_pf_nt<G, RM, 256, 6>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w,
                                        tile0);
    else
        launch_reg_pf_nt<G, RM, 256>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w,
                                     tile0);
}

// The register x pass over all ntiles tiles of the plane.
template <int G, unsigned RM>
static void launch_reg_pf(dim3 grid, size_t lds, hipStream_t st, const float2* spec, const float2* tgt,
                          float2* out, const FftDesc& fd, int n, int nhz, int nb, int C, int T, int ntiles,
                          unsigned grp_mask, int lo, int w, bool use_static) {
    // compile-time tile plans: the full 8-column tiles (4-column above grid
    // edge 256), with the same threads and values per thread as the runtime
    // form; a last partial tile (when n (n / 2 + 1) is not a multiple of the
    // tile width) runs the runtime form
    const int full = n * nhz / T;
    auto launch_static = [&](auto sp, auto width) {
        using SP = decltype(sp);
        constexpr int TW = decltype(width)::value;
        using TL = typename SP::template Tile<TW>;
        constexpr int E = SP::kN * TW;
        constexpr int NT = E <= 64 * kXcorrPF ? 64 : (E <= 128 * kXcorrPF ? 128 : 256);
        constexpr int PF = (NT == 256 && E <= 256 * 6) ? 6 : kXcorrPF;
        static_assert(E <= NT * PF, "tile exceeds the workgroup's values");
        launch_reg_pf_nt<G, SP::kMask, NT, PF, TL>(xcorr_grid(full, nb), lds, st, spec, tgt, out, fd, n, nhz, nb, C,
                                                   T, full, grp_mask, lo, w, 0);
    };
    auto tiles8 = [&](auto sp) { launch_static(sp, std::integral_constant<int, 8>{}); };
    auto tiles4 = [&](auto sp) { launch_static(sp, std::integral_constant<int, 4>{}); };
    bool hit = false;
    if (use_static && full > 0 && T == 8)
        hit = with_static_plan(fd, StaticLengths{}, tiles8) || with_static_plan(fd, SmallLengths{}, tiles8) ||
              with_static_plan(fd, MediumLengths{}, tiles8) || with_static_plan(fd, TileLengths{}, tiles8);
    else if (use_static && full > 0 && T == 4 && n > 256)
        hit = with_static_plan(fd, LargeLengths{}, tiles4);
    if (hit) {
        if (full < ntiles)
            launch_reg_pf_dyn<G, RM>(xcorr_grid(ntiles - full, nb), lds, st, spec, tgt, out, fd, n, nhz, nb, C, T,
                                     ntiles - full, grp_mask, lo, w, full);
        return;
    }
    launch_reg_pf_dyn<G, RM>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w, 0);
}

int launch_xcorr(const float2* spec, const float2* tgt, float2* out, const FftDesc& fd, int n,
                 int nb, int C, int G, unsigned grp_mask, int lo, int w, hipStream_t st,
                 int xvariant = 0, int xT = 0, bool use_static = true) {
    const int nhz = n / 2 + 1;
    if (tgt != nullptr && xvariant == 3) {
        const int T = xT > 0 ? xT : (n * 8 <= 256 * kXcorrPF ? 8 : 4);
        VX_REQUIRE(n * T <= 256 * kXcorrPF, VX_ERR_UNSUPPORTED,
                   "x-pass variant 3 needs grid edge x tile <= %d", 256 * kXcorrPF);
        const size_t lds = ((size_t)2 * n * T + fd.ntw) * sizeof(float2);
        VX_REQUIRE(lds <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d too large for the x pass", n);
        const int ntiles = (n * nhz + T - 1) / T;
        dim3 grid((unsigned)(((ntiles + 7) / 8) * 8 * nb));
        const unsigned rm = radix_mask(fd);
        if (G == 1) {
            if (rm == kRadix57) launch_reg_pf<1, kRadix57>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w, use_static);
            else                launch_reg_pf<1, kAllRadices>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w, use_static);
        } else {
            if (rm == kRadix57) launch_reg_pf<2, kRadix57>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w, use_static);
            else                launch_reg_pf<2, kAllRadices>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w, use_static);
        }
        VX_LAUNCH_CHECK();
        return VX_OK;
    }
    if (tgt == nullptr) {
        // target preparation: forward x FFT + conj per channel (once per search)
        const int T = xcorr_T(n);
        VX_REQUIRE(n * T <= 256 * 16, VX_ERR_UNSUPPORTED, "grid edge %d too large for the x pass", n);
        const size_t lds = ((size_t)n * T + fd.ntw) * sizeof(float2);
        dim3 grid((n * nhz + T - 1) / T, nb);
        hipLaunchKernelGGL((k_xcorr<16, 1>), grid, dim3(256), lds, st, spec, tgt, out, fd, n, nhz, C, T, grp_mask);
        VX_LAUNCH_CHECK();
        return VX_OK;
    }
    if (xvariant == 1) {
        // default tile: 8 columns, 4 above grid edge 256 (the (2 + G) tile
        // buffers then stay near 32 KB, four workgroups per CU; measured at 343)
        int T = xT > 0 ? xT : (n <= 256 ? 8 : 4);
        while (T > 1 && ((size_t)(2 + G) * n * T + fd.ntw) * sizeof(float2) > kMaxLds) T >>= 1;
        const size_t lds = ((size_t)(2 + G) * n * T + fd.ntw) * sizeof(float2);
        VX_REQUIRE(lds <= kMaxLds, VX_ERR_UNSUPPORTED, "grid edge %d too large for the x pass", n);
        const int ntiles = (n * nhz + T - 1) / T;
        dim3 grid((unsigned)(((ntiles + 7) / 8) * 8 * nb));
        if (n * T <= 256 * kXcorrPF) {
            const unsigned rm = radix_mask(fd);
            if (G == 1) {
                if (rm == kRadix57) launch_pf<1, kRadix57>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
                else                launch_pf<1, kAllRadices>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
            } else {
                if (rm == kRadix57) launch_pf<2, kRadix57>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
                else                launch_pf<2, kAllRadices>(grid, lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
            }
            VX_LAUNCH_CHECK();
            return VX_OK;
        }
        if (lds > 65536) {
            if (G == 1) VX_HIP(hipFuncSetAttribute((const void*)k_xcorr_lds<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            else        VX_HIP(hipFuncSetAttribute((const void*)k_xcorr_lds<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        }
        if (G == 1) hipLaunchKernelGGL((k_xcorr_lds<1>), grid, dim3(256), lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
        else        hipLaunchKernelGGL((k_xcorr_lds<2>), grid, dim3(256), lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
        VX_LAUNCH_CHECK();
        return VX_OK;
    }
    const int T = xcorr_pp_T(n);
    VX_REQUIRE(n * T <= 256 * kXcorrMaxE, VX_ERR_UNSUPPORTED, "grid edge %d too large for the x pass", n);
    const size_t lds = (2 * (size_t)n * T + fd.ntw) * sizeof(float2);
    const int ntiles = (n * nhz + T - 1) / T;
    const int groups8 = (ntiles + 7) / 8;
    dim3 grid((unsigned)(groups8 * 8 * nb));
    if (G == 1) hipLaunchKernelGGL((k_xcorr_pp<kXcorrMaxE, 1>), grid, dim3(256), lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
    else        hipLaunchKernelGGL((k_xcorr_pp<kXcorrMaxE, 2>), grid, dim3(256), lds, st, spec, tgt, out, fd, n, nhz, nb, C, T, ntiles, grp_mask, lo, w);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

int launch_gather(int kind, const float4* svec, const float2* sscal, const float4* sprob,
                  const double* sdat, const float* rot, int n, int nb, int fma, float* q,
                  double* qd, unsigned long long* keys, int nkeys, int lo, int w, int slo, int sw,
                  hipStream_t st) {
    const int nzb = (w + 63) / 64;
    dim3 grid(nzb * ((w + 3) / 4), w, nb);
    switch (kind) {
        case KIND_VEC:    hipLaunchKernelGGL(k_gather<KIND_VEC>, grid, dim3(256), 0, st, svec, sscal, sprob, sdat, rot, n, nzb, fma, q, qd, keys, nkeys, lo, w, slo, sw); break;
        case KIND_SCALAR: hipLaunchKernelGGL(k_gather<KIND_SCALAR>, grid, dim3(256), 0, st, svec, sscal, sprob, sdat, rot, n, nzb, fma, q, qd, keys, nkeys, lo, w, slo, sw); break;
        case KIND_PROB:   hipLaunchKernelGGL(k_gather<KIND_PROB>, grid, dim3(256), 0, st, svec, sscal, sprob, sdat, rot, n, nzb, fma, q, qd, keys, nkeys, lo, w, slo, sw); break;
        default:          hipLaunchKernelGGL(k_gather<KIND_DATA64>, grid, dim3(256), 0, st, svec, sscal, sprob, sdat, rot, n, nzb, fma, q, qd, keys, nkeys, lo, w, slo, sw); break;
    }
    VX_LAUNCH_CHECK();
    return VX_OK;
}

}  // namespace vx

// ---------------------------------------------------------------------------
// Plan C-ABI
// ---------------------------------------------------------------------------

using namespace vx;

struct vx_plan {
    int n, nhz, mode, batch, flags;
    int C;        // query channels
    int G;        // correlation groups
    int nkeys;    // keys per rotation (G, +1 for the prob mixed score)
    int kind;     // gather kind
    FftDesc fd;
    float2* twid;
    float4* svec;
    float2* sscal;
    float4* sprob;
    double* sdat;
    float* q;
    double* qd;
    float2* spec;
    float2* sum;
    float2* tgt;
    unsigned long long* keys;
    float* real_out;
    float* rot_stage;
    bool have_target, have_search, mix;
    bool q_dirty;     // q / qd hold non-zero values outside the support box
    int src_lo, src_w; // box of the compact gather source (svec / sscal / sprob)
    int lo, w;        // support box [lo, lo + w)^3 of the rotated query
    int xvariant;     // x-pass kernel: 0 register FFT, 1 LDS FFT, 2 MFMA DFT, 3 register-prefetch FFT,
                      // 4 MFMA radix split (N = N1 N2, length-N1 DFTs on split-f16 MFMA)
    int xT;           // x-pass tile width (columns), 0 = default
    int pencil;       // 1: separate z and y passes (slab does not fit LDS, or forced)
    int pin_w;        // > 0: support box fixed to [pin_lo, pin_lo + pin_w) (graph replay across pairs)
    int pin_lo;
    int slab_nt;      // preferred threads per slab workgroup (0 = default)
    int nstreams;     // 2: alternate batches over two workspaces and streams
    int static_plans; // 1: compile-time stage plans for the grid edges that have one
    int slab_fused;   // 1: fused slab kernels on compile-time plans (k_fwd_zy_fused, k_inv_yz_fused)
    int pencil_chunks; // line-pair chunks per workgroup of the fused inverse z pass (0 = automatic)
    float* q2;        // second workspace (nstreams == 2)
    double* qd2;
    float2* spec2;
    float2* sum2;
    unsigned long long* keys2;
    float* real_out2;
    hipStream_t st2;
    hipEvent_t ev_fork, ev_join;
    unsigned short* dft;     // MFMA x pass: split bf16 F (hi, lo) then G (hi, lo), fragment order
    size_t dft_elems[2];     // elements of one F matrix, one G matrix
    int dft_lo, dft_w, dft_kp; // box the F matrices were built for (dft_w = 0: none)
    float2* mr_w;            // variant 4: W_N^j = exp(-2 pi i j / N), j < N
    uint32_t* mr_a;          // variant 4: A operands of the support box (xcorr_mr_tables)
    int mr_lo, mr_bw;        // box they were built for (mr_bw = 0: none)
    double alpha, vstd, vave, pstd, pave;
    size_t bytes;
    // FP64 re-scoring workspace (vx_rescore_f64), grown on demand
    float* rs_maps;        // [batch][G][n^3] float32 correlation maps of a batch
    size_t rs_maps_cap;
    float* rs_best;        // [rows][gout] float32 maxima
    int32_t* rs_trans;     // [rows][gout][3]
    int* rs_ncand;         // [rows][gout]
    int* rs_cand;          // [rows][gout][max_cand]
    double* rs_partial;    // [rows][gout][max_cand][split][2]
    size_t rs_rows_cap;    // rows x gout x max_cand the buffers above hold
    int rs_rows_alloc, rs_mc_alloc;
};

static void plan_free(vx_plan* p) {
    if (!p) return;
    void* ptrs[] = {p->twid, p->svec, p->sscal, p->sprob, p->sdat, p->q, p->qd, p->spec,
                    p->sum, p->tgt, p->keys, p->real_out, p->rot_stage, p->dft, p->mr_w, p->mr_a,
                    p->q2, p->qd2, p->spec2, p->sum2, p->keys2, p->real_out2,
                    p->rs_maps, p->rs_best, p->rs_trans, p->rs_ncand, p->rs_cand, p->rs_partial};
    for (void* x : ptrs)
        if (x) (void)hipFree(x);
    if (p->st2) (void)hipStreamDestroy(p->st2);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    delete p;
}

#define PLAN_ALLOC(field, count)                                                        \
    do {                                                                                \
        const size_t b_ = (size_t)(count) * sizeof(*p->field);                          \
        if (hipMalloc((void**)&p->field, b_) != hipSuccess) {                           \
            plan_free(p);                                                               \
            (void)hipGetLastError();                                                    \
            set_error("device allocation of %zu bytes for " #field " failed", b_);      \
            return VX_ERR_NOMEM;                                                        \
        }                                                                               \
        p->bytes += b_;                                                                 \
    } while (0)

extern "C" int vx_plan_create(int n, int mode, int batch, int flags, vx_plan** out) {
    VX_REQUIRE(out != nullptr, VX_ERR_ARG, "out is NULL");
    *out = nullptr;
    VX_REQUIRE(n >= 2 && n <= 1024, VX_ERR_ARG, "grid edge %d out of range", n);
    VX_REQUIRE(mode >= VX_MODE_VEC && mode <= VX_MODE_PROB, VX_ERR_ARG, "bad mode %d", mode);
    VX_REQUIRE(batch >= 1 && batch <= 4096, VX_ERR_ARG, "bad batch %d", batch);
    const int nhz = n / 2 + 1;
    vx_plan* p = new (std::nothrow) vx_plan();
    VX_REQUIRE(p != nullptr, VX_ERR_NOMEM, "host allocation failed");
    memset(p, 0, sizeof(*p));
    p->n = n; p->nhz = nhz; p->mode = mode; p->batch = batch; p->flags = flags;
    p->lo = 0; p->w = n; p->q_dirty = true;
    // register-accumulating prefetch x pass (the autotuner's usual pick on
    // MI355X) with 8-column tiles, 4-column above grid edge 256 (4-6 % faster
    // than the LDS x pass there, N = 300..360), else the LDS one
    p->xvariant = n * 4 <= 256 * kXcorrPF ? 3 : 1;
    p->xT = 0;
    p->nstreams = 1;
    p->static_plans = 1;
    p->slab_fused = 1;
    p->pencil_chunks = 0;
    switch (mode) {
        case VX_MODE_VEC:  p->C = 3; p->G = 1; p->nkeys = 1; p->kind = KIND_VEC; break;
        case VX_MODE_PROB: p->C = 7; p->G = 2; p->nkeys = 3; p->kind = KIND_PROB; break;
        case VX_MODE_LAPLACIAN: p->C = 1; p->G = 1; p->nkeys = 1; p->kind = KIND_DATA64; break;
        default:           p->C = 1; p->G = 1; p->nkeys = 1; p->kind = KIND_SCALAR; break;
    }
    int rc = fft_desc_create(n, &p->fd, &p->twid);
    if (rc != VX_OK) { plan_free(p); return rc; }
    {
        int nthr, maxv;
        slab_shape(n, nhz, &nthr, &maxv);
        p->pencil = (maxv == 0 || ((size_t)n * nhz + p->fd.ntw) * sizeof(float2) > kMaxLds) ? 1 : 0;
    }
    const size_t n3 = (size_t)n * n * n;
    const size_t nh = (size_t)n * n * nhz;
    const int B = batch;
    if (p->kind == KIND_VEC || p->kind == KIND_PROB) PLAN_ALLOC(svec, n3);
    if (p->kind == KIND_PROB) PLAN_ALLOC(sprob, n3);
    if (p->kind == KIND_SCALAR) PLAN_ALLOC(sscal, n3);
    if (p->kind == KIND_DATA64) { PLAN_ALLOC(sdat, n3); PLAN_ALLOC(qd, (size_t)B * n3); }
    PLAN_ALLOC(q, (size_t)std::max(B, 1) * p->C * n3);
    PLAN_ALLOC(spec, (size_t)B * p->C * nh);
    PLAN_ALLOC(sum, (size_t)B * p->G * nh);
    PLAN_ALLOC(tgt, (size_t)p->C * nh);
    PLAN_ALLOC(keys, (size_t)B * p->nkeys);
    PLAN_ALLOC(rot_stage, (size_t)9 * B);
    if (mode == VX_MODE_PROB) PLAN_ALLOC(real_out, (size_t)B * 2 * n3);
    *out = p;
    return VX_OK;
}

extern "C" int vx_plan_destroy(vx_plan* p) {
    if (p) {
        (void)hipDeviceSynchronize();
        plan_free(p);
    }
    return VX_OK;
}

extern "C" int vx_plan_info(const vx_plan* p, int64_t* info) {
    VX_REQUIRE(p && info, VX_ERR_ARG, "NULL argument");
    info[0] = p->n; info[1] = p->mode; info[2] = p->batch;
    info[3] = p->mix ? 3 : p->nkeys == 3 ? 2 : p->G;
    info[4] = p->C; info[5] = (int64_t)p->bytes; info[6] = p->lo; info[7] = p->w;
    return VX_OK;
}

// (Re)build the radix-split MFMA x pass's A operands for the plan's support box.
static int ensure_mr(vx_plan* p) {
    if (p->xvariant != 4 || (p->mr_a && p->mr_lo == p->lo && p->mr_bw == p->w)) return VX_OK;
    std::vector<uint32_t> tab;
    int rc = xcorr_mr_tables(p->n, p->lo, p->w, tab);
    if (rc != VX_OK) return rc;
    constexpr size_t kCap = (size_t)4 * 4 * 2 * 64 * 4;   // KSF + KSI <= 4
    if (!p->mr_a) {
        if (hipMalloc((void**)&p->mr_a, kCap * sizeof(uint32_t)) != hipSuccess) {
            (void)hipGetLastError();
            p->mr_a = nullptr;
            set_error("device allocation of the MFMA x-pass operands failed");
            return VX_ERR_NOMEM;
        }
        p->bytes += kCap * sizeof(uint32_t);
    }
    VX_HIP(hipDeviceSynchronize());   // no launch in flight may still read the old operands
    VX_HIP(hipMemcpy(p->mr_a, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    p->mr_lo = p->lo;
    p->mr_bw = p->w;
    return VX_OK;
}

// (Re)build the MFMA x pass's DFT matrices for the plan's support box.
static int ensure_dft(vx_plan* p) {
    const int rc = ensure_mr(p);
    if (rc != VX_OK) return rc;
    if (p->xvariant != 2 || (p->dft && p->dft_lo == p->lo && p->dft_w == p->w)) return VX_OK;
    const int n = p->n, NP = 16 * ((n + 15) / 16);
    const int KP = (p->w + 15) / 16 * 16;
    std::vector<unsigned short> fh, fl, gh, gl;
    build_dft_matrix(n, NP, KP, p->lo, p->w, true, fh, fl);
    build_dft_matrix(n, NP, NP, 0, n, false, gh, gl);
    const size_t fe = (size_t)2 * NP * 2 * NP;   // capacity: KP <= NP
    if (!p->dft) {
        if (hipMalloc((void**)&p->dft, 4 * fe * sizeof(unsigned short)) != hipSuccess) {
            (void)hipGetLastError();
            p->dft = nullptr;
            set_error("device allocation of the DFT matrices failed");
            return VX_ERR_NOMEM;
        }
        p->bytes += 4 * fe * sizeof(unsigned short);
    }
    p->dft_elems[0] = fe;
    p->dft_elems[1] = fe;
    VX_HIP(hipDeviceSynchronize());   // no launch in flight may still read the old matrices
    unsigned short* d = p->dft;
    VX_HIP(hipMemcpy(d, fh.data(), fh.size() * 2, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(d + fe, fl.data(), fl.size() * 2, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(d + 2 * fe, gh.data(), gh.size() * 2, hipMemcpyHostToDevice));
    VX_HIP(hipMemcpy(d + 3 * fe, gl.data(), gl.size() * 2, hipMemcpyHostToDevice));
    p->dft_lo = p->lo;
    p->dft_w = p->w;
    p->dft_kp = KP;
    return VX_OK;
}

namespace vx {
// Largest squared distance from the rotation centre n/2 of a voxel whose
// density (or, in prob mode, any probability channel) is non-zero.
__global__ void __launch_bounds__(256)
k_support_r2(const double* __restrict__ data, const float* __restrict__ prob, int n,
             unsigned long long* __restrict__ out) {
    const size_t n3 = (size_t)n * n * n;
    const double c = 0.5 * n;
    double best = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        bool nz = data[i] != 0.0;
        if (prob) nz = nz || prob[i] != 0.f || prob[n3 + i] != 0.f || prob[2 * n3 + i] != 0.f || prob[3 * n3 + i] != 0.f;
        if (nz) {
            const int z = (int)(i % n), y = (int)((i / n) % n), x = (int)(i / ((size_t)n * n));
            const double dx = x - c, dy = y - c, dz = z - c;
            best = fmax(best, dx * dx + dy * dy + dz * dz);
        }
    }
    for (int off = 32; off > 0; off >>= 1) best = fmax(best, __shfl_xor(best, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(best));
}
}  // namespace vx

extern "C" int vx_set_search(vx_plan* p, const double* sdata, const float* svec, const float* sprob,
                             double t_ave, vx_stream stream) {
    VX_REQUIRE(p && sdata, VX_ERR_ARG, "NULL plan or data");
    hipStream_t st = (hipStream_t)stream;
    const size_t n3 = (size_t)p->n * p->n * p->n;
    if (p->kind == KIND_VEC || p->kind == KIND_PROB) VX_REQUIRE(svec, VX_ERR_ARG, "vec required");
    if (p->kind == KIND_PROB) VX_REQUIRE(sprob, VX_ERR_ARG, "prob channels required");
    if (p->kind == KIND_DATA64)
        VX_HIP(hipMemcpyAsync(p->sdat, sdata, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    // Support box of every rotation of the query: a destination voxel d takes
    // source voxel floor(F (i - c) + c) with |i - d| <= 1 per axis (odd-N
    // fold), so a non-zero value lands only within radius R_s + 2 sqrt(3) of c,
    // R_s = the largest centre distance of a non-zero source voxel.
    if (p->flags & VX_FLAG_NO_PRUNE) {
        p->lo = 0;
        p->w = p->n;
    } else {
        unsigned long long* d_r2 = p->keys;   // scratch: keys are rewritten by every batch
        VX_HIP(hipMemsetAsync(d_r2, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_support_r2, dim3(512), dim3(256), 0, st, sdata,
                           p->kind == KIND_PROB ? sprob : nullptr, p->n, d_r2);
        VX_LAUNCH_CHECK();
        unsigned long long h = 0;
        VX_HIP(hipMemcpyAsync(&h, d_r2, sizeof(h), hipMemcpyDeviceToHost, st));
        VX_HIP(hipStreamSynchronize(st));
        double r2;
        memcpy(&r2, &h, sizeof(r2));
        const double c = 0.5 * p->n;
        const double r = sqrt(r2) + 2.0 * sqrt(3.0) + 1.0;
        int lo = (int)floor(c - r), hi = (int)ceil(c + r);
        if (lo < 0) lo = 0;
        if (hi > p->n - 1) hi = p->n - 1;
        if (p->pin_w > 0) {
            const int plo = p->pin_lo, phi = plo + p->pin_w - 1;
            VX_REQUIRE(lo >= plo && hi <= phi, VX_ERR_STATE,
                       "query support [%d, %d] exceeds the pinned box [%d, %d]", lo, hi, plo, phi);
            lo = plo;
            hi = phi;
        }
        // planes hold the previous box's values outside the new box: clear them
        if (lo != p->lo || hi - lo + 1 != p->w) p->q_dirty = true;
        p->lo = lo;
        p->w = hi - lo + 1;
    }
    // the gather's source: the box of the query, compact (every non-zero voxel
    // lies inside it)
    if (p->kind != KIND_DATA64) {
        const size_t nb3 = (size_t)p->w * p->w * p->w;
        hipLaunchKernelGGL(k_build_src, dim3((unsigned)((nb3 + 255) / 256)), dim3(256), 0, st, sdata, svec, sprob,
                           n3, p->mode, t_ave, p->svec, p->sscal, p->sprob, p->n, p->lo, p->w);
        VX_LAUNCH_CHECK();
    }
    p->src_lo = p->lo;
    p->src_w = p->w;
    p->have_search = true;
    return ensure_dft(p);
}

extern "C" int vx_set_target(vx_plan* p, const double* tdata, const float* tvec, const float* tprob,
                             double t_ave, vx_stream stream) {
    VX_REQUIRE(p && tdata, VX_ERR_ARG, "NULL plan or data");
    if (p->mode == VX_MODE_VEC || p->mode == VX_MODE_PROB) VX_REQUIRE(tvec, VX_ERR_ARG, "vec required");
    if (p->mode == VX_MODE_PROB) VX_REQUIRE(tprob, VX_ERR_ARG, "prob channels required");
    hipStream_t st = (hipStream_t)stream;
    const int n = p->n;
    const size_t n3 = (size_t)n * n * n;
    double* lap = nullptr;
    if (p->mode == VX_MODE_LAPLACIAN) {
        lap = p->qd;   // batch >= 1 float64 planes available
        int rc = vx_laplacian(n, tdata, lap, stream);
        if (rc != VX_OK) return rc;
    }
    hipLaunchKernelGGL(k_target_chan, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st,
                       tdata, tvec, tprob, lap, n3, p->mode, t_ave, p->q);
    VX_LAUNCH_CHECK();
    // C planes of batch slot 0 -> spectra (x-slab pass), then x pass with conj.
    // spec holds B*C planes >= C planes.
    int rc = p->pencil ? launch_fwd_pencil(p->q, p->spec, p->fd, n, p->C, 0, n, st, p->static_plans != 0, p->slab_fused != 0)
                       : launch_fwd_zy(p->q, p->spec, p->fd, n, p->C, 0, n, st, 0, p->static_plans != 0,
                                       p->slab_fused != 0);
    if (rc != VX_OK) return rc;
    // treat the C channels as "batch" 1 with C channels, forward-conj mode
    rc = launch_xcorr(p->spec, nullptr, p->tgt, p->fd, n, 1, p->C, 1, 0u, 0, n, st);
    if (rc != VX_OK) return rc;
    // the target channels used the query planes: clear them for the gather,
    // which writes only inside the support box
    VX_HIP(hipMemsetAsync(p->q, 0, (size_t)p->batch * p->C * n3 * sizeof(float), st));
    if (p->qd) VX_HIP(hipMemsetAsync(p->qd, 0, (size_t)p->batch * n3 * sizeof(double), st));
    p->have_target = true;
    p->q_dirty = false;
    return VX_OK;
}

// Second workspace and stream for two batches in flight (VX_TUNE_STREAMS 2).
static int alloc_second_workspace(vx_plan* p) {
    const size_t n3 = (size_t)p->n * p->n * p->n;
    const size_t nh = (size_t)p->n * p->n * p->nhz;
    const int B = p->batch;
    struct Req { void** ptr; size_t bytes; };
    Req reqs[] = {{(void**)&p->q2, (size_t)B * p->C * n3 * sizeof(float)},
                  {(void**)&p->qd2, p->kind == KIND_DATA64 ? (size_t)B * n3 * sizeof(double) : 0},
                  {(void**)&p->spec2, (size_t)B * p->C * nh * sizeof(float2)},
                  {(void**)&p->sum2, (size_t)B * p->G * nh * sizeof(float2)},
                  {(void**)&p->keys2, (size_t)B * p->nkeys * sizeof(unsigned long long)},
                  {(void**)&p->real_out2, p->mode == VX_MODE_PROB ? (size_t)B * 2 * n3 * sizeof(float) : 0}};
    for (auto& r : reqs) {
        if (r.bytes == 0) continue;
        if (hipMalloc(r.ptr, r.bytes) != hipSuccess) {
            (void)hipGetLastError();
            *r.ptr = nullptr;
            set_error("device allocation of %zu bytes for the second workspace failed", r.bytes);
            return VX_ERR_NOMEM;
        }
        p->bytes += r.bytes;
    }
    VX_HIP(hipMemset(p->q2, 0, (size_t)B * p->C * n3 * sizeof(float)));
    if (p->qd2) VX_HIP(hipMemset(p->qd2, 0, (size_t)B * n3 * sizeof(double)));
    VX_HIP(hipStreamCreateWithFlags(&p->st2, hipStreamNonBlocking));
    VX_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    VX_HIP(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
    return VX_OK;
}

extern "C" int vx_plan_tune(vx_plan* p, int key, int value) {
    VX_REQUIRE(p, VX_ERR_ARG, "NULL plan");
    switch (key) {
        case VX_TUNE_XPASS_VARIANT:
            VX_REQUIRE(value >= 0 && value <= 4, VX_ERR_ARG, "x-pass variant must be 0, 1, 2, 3 or 4");
            if (value == 4) {
                int n1, n2;
                VX_REQUIRE(p->G == 1, VX_ERR_UNSUPPORTED, "the MFMA radix-split x pass has one correlation group");
                VX_REQUIRE(xcorr_mr_factor(p->n, &n1, &n2), VX_ERR_UNSUPPORTED,
                           "grid edge %d has no factor N1 * N2 (N1 <= 32, N2 in 2, 3, 4, 5, 7, 8)", p->n);
                if (!p->mr_w) {
                    std::vector<float2> wt(p->n);
                    for (int j = 0; j < p->n; ++j) {
                        const double a = -2.0 * M_PI * (double)j / p->n;
                        wt[j] = make_float2((float)cos(a), (float)sin(a));
                    }
                    if (hipMalloc((void**)&p->mr_w, p->n * sizeof(float2)) != hipSuccess) {
                        (void)hipGetLastError();
                        p->mr_w = nullptr;
                        set_error("device allocation of the x-pass twiddles failed");
                        return VX_ERR_NOMEM;
                    }
                    VX_HIP(hipMemcpy(p->mr_w, wt.data(), p->n * sizeof(float2), hipMemcpyHostToDevice));
                    p->bytes += p->n * sizeof(float2);
                }
            }
            if (value == 2) {
                VX_REQUIRE(p->G == 1, VX_ERR_UNSUPPORTED, "the MFMA x pass has one correlation group");
                VX_REQUIRE((p->n + 15) / 16 <= kMfmaMaxNPT &&
                           xcorr_mfma_lds(16 * ((p->n + 15) / 16), 16 * ((p->n + 15) / 16), 64) <= kMaxLds,
                           VX_ERR_UNSUPPORTED, "grid edge %d too large for the MFMA x pass", p->n);
            }
            p->xvariant = value; break;
        case VX_TUNE_XPASS_TILE:
            VX_REQUIRE(value >= 0 && value <= 64, VX_ERR_ARG, "x-pass tile must be in [0, 64]");
            p->xT = value; break;
        case VX_TUNE_PIN_BOX_LO:
            VX_REQUIRE(value >= 0 && value < p->n, VX_ERR_ARG, "pinned box start must be in [0, n)");
            p->pin_lo = value;
            break;
        case VX_TUNE_PIN_BOX_WIDTH:
            VX_REQUIRE(value >= 0 && p->pin_lo + value <= p->n, VX_ERR_ARG, "pinned box exceeds the grid");
            p->pin_w = value;
            if (value > 0 && (p->lo != p->pin_lo || p->w != value)) {
                p->lo = p->pin_lo;
                p->w = value;
                p->q_dirty = true;
            }
            break;
        case VX_TUNE_STREAMS:
            VX_REQUIRE(value == 1 || value == 2, VX_ERR_ARG, "streams must be 1 or 2");
            if (value == 2 && !p->q2) {
                int rc = alloc_second_workspace(p);
                if (rc != VX_OK) return rc;
            }
            p->nstreams = value;
            break;
        case VX_TUNE_STATIC_PLANS:
            VX_REQUIRE(value == 0 || value == 1, VX_ERR_ARG, "static plans must be 0 or 1");
            p->static_plans = value;
            break;
        case VX_TUNE_SLAB_FUSED:
            VX_REQUIRE(value == 0 || value == 1, VX_ERR_ARG, "slab fusion must be 0 or 1");
            p->slab_fused = value;
            break;
        case VX_TUNE_PENCIL_CHUNKS:
            VX_REQUIRE(value >= 0, VX_ERR_ARG, "pencil chunks must be >= 0");
            p->pencil_chunks = value;
            break;
        case VX_TUNE_SLAB_THREADS:
            VX_REQUIRE(value == 0 || value == 512 || value == 1024, VX_ERR_ARG,
                       "slab threads must be 0, 512 or 1024");
            p->slab_nt = value;
            break;
        case VX_TUNE_RADIX_ORDER: {
            VX_REQUIRE(value >= 0 && value <= 3, VX_ERR_ARG, "radix order must be 0, 1, 2 or 3");
            FftDesc nd;
            float2* ntw = nullptr;
            int rc = fft_desc_create(p->n, &nd, &ntw, value);
            if (rc != VX_OK) return rc;
            VX_HIP(hipDeviceSynchronize());   // no launch in flight may still read the old stage table
            (void)hipFree(p->twid);
            p->fd = nd;
            p->twid = ntw;
            break;
        }
        case VX_TUNE_FORCE_PENCIL:
            if (value) p->pencil = 1;
            else {
                int nthr, maxv;
                slab_shape(p->n, p->nhz, &nthr, &maxv);
                VX_REQUIRE(maxv > 0 && ((size_t)p->n * p->nhz + p->fd.ntw) * sizeof(float2) <= kMaxLds,
                           VX_ERR_UNSUPPORTED, "grid edge %d needs the pencil passes", p->n);
                p->pencil = 0;
            }
            break;
        default:
            VX_REQUIRE(false, VX_ERR_ARG, "unknown tuning key %d", key);
    }
    return ensure_dft(p);
}

extern "C" int vx_set_prob_stats(vx_plan* p, int enable, double alpha, double vstd, double vave,
                                 double pstd, double pave) {
    VX_REQUIRE(p, VX_ERR_ARG, "NULL plan");
    VX_REQUIRE(p->mode == VX_MODE_PROB, VX_ERR_STATE, "plan is not in prob mode");
    p->mix = enable != 0;
    p->alpha = alpha; p->vstd = vstd; p->vave = vave; p->pstd = pstd; p->pave = pave;
    return VX_OK;
}

// Stage boundaries for vx_score_rows_profile: 0 gather, 1 fwd_zy, 2 xcorr,
// 3 inv_yz, 4 mixed, 5 finalize.
constexpr int kStages = 6;

// One batch workspace: query planes, spectra, correlation spectrum, keys.
struct Workspace {
    float* q;
    double* qd;
    float2* spec;
    float2* sum;
    unsigned long long* keys;
    float* real_out;
};

// Score rows [r0, r0 + nb) in workspace ws on stream st; ev (7 events) marks
// the stage boundaries for the profiler.
static int score_batch(vx_plan* p, const Workspace& ws, const float* rot, int r0, int nb, float* best,
                       int32_t* trans, hipStream_t st, hipEvent_t* ev, float* maps = nullptr) {
    const int n = p->n;
    const size_t n3 = (size_t)n * n * n;
    const int nk = p->nkeys;
    const int nout = (p->mode == VX_MODE_PROB) ? (p->mix ? 3 : 2) : 1;
    const unsigned grp_mask = (p->mode == VX_MODE_PROB) ? 0x78u : 0u;
#define VX_MARK(i) do { if (ev) VX_HIP(hipEventRecord(ev[i], st)); } while (0)
    VX_MARK(0);
    int rc = launch_gather(p->kind, p->svec, p->sscal, p->sprob, p->sdat, rot + (size_t)r0 * 9, n, nb,
                           (p->flags & VX_FLAG_NO_FMA) ? 0 : 1, ws.q, ws.qd, ws.keys, nk, p->lo, p->w, p->src_lo,
                           p->src_w, st);
    if (rc != VX_OK) return rc;
    if (p->kind == KIND_DATA64) {
        hipLaunchKernelGGL(k_lap_to_q, dim3((unsigned)((n3 + 255) / 256), nb), dim3(256), 0, st, ws.qd, ws.q, n);
        VX_LAUNCH_CHECK();
    }
    VX_MARK(1);
    rc = p->pencil ? launch_fwd_pencil(ws.q, ws.spec, p->fd, n, nb * p->C, p->lo, p->w, st, p->static_plans != 0, p->slab_fused != 0)
                   : launch_fwd_zy(ws.q, ws.spec, p->fd, n, nb * p->C, p->lo, p->w, st, p->slab_nt,
                                   p->static_plans != 0, p->slab_fused != 0);
    if (rc != VX_OK) return rc;
    VX_MARK(2);
    if (p->xvariant == 2) {
        VX_REQUIRE(p->dft && p->dft_lo == p->lo && p->dft_w == p->w, VX_ERR_STATE,
                   "MFMA x-pass matrices do not match the support box");
        const unsigned short* mats[4] = {p->dft, p->dft + p->dft_elems[0], p->dft + 2 * p->dft_elems[0],
                                         p->dft + 2 * p->dft_elems[0] + p->dft_elems[1]};
        rc = launch_xcorr_mfma(ws.spec, p->tgt, ws.sum, mats, n, nb, p->C, p->dft_kp, p->lo, p->w, p->xT, st);
    } else if (p->xvariant == 4) {
        VX_REQUIRE(p->mr_a && p->mr_lo == p->lo && p->mr_bw == p->w, VX_ERR_STATE,
                   "MFMA x-pass operands do not match the support box");
        rc = launch_xcorr_mr(ws.spec, p->tgt, ws.sum, p->mr_w, p->mr_a, n, nb, p->C, p->lo, p->w, st);
    } else {
        rc = launch_xcorr(ws.spec, p->tgt, ws.sum, p->fd, n, nb, p->C, p->G, grp_mask, p->lo, p->w, st,
                          p->xvariant, p->xT, p->static_plans != 0);
    }
    if (rc != VX_OK) return rc;
    VX_MARK(3);
    // maps != nullptr: every group's real correlation map is kept (FP64
    // re-scoring picks its candidates there); prob mixing reads them too
    float* ro = maps ? maps : (p->mode == VX_MODE_PROB && p->mix) ? ws.real_out : nullptr;
    rc = p->pencil ? launch_inv_pencil(ws.sum, p->fd, n, nb * p->G, p->G, ws.keys, nk, ro, st, p->static_plans != 0, p->slab_fused != 0,
                                       p->pencil_chunks)
                   : launch_inv_yz(ws.sum, p->fd, n, nb * p->G, p->G, ws.keys, nk, ro, st, p->slab_nt,
                                   p->static_plans != 0, p->slab_fused != 0);
    if (rc != VX_OK) return rc;
    VX_MARK(4);
    if (p->mode == VX_MODE_PROB && p->mix) {
        const float wa = (float)(1.0 - p->alpha), wb = (float)p->alpha;
        unsigned gx = (unsigned)std::min<size_t>((n3 + 255) / 256, 512);
        hipLaunchKernelGGL(k_mixed, dim3(gx, nb), dim3(256), 0, st, ro, n3,
                           (float)p->vave, (float)p->vstd, (float)p->pave, (float)p->pstd, wa, wb,
                           ws.keys, nk);
        VX_LAUNCH_CHECK();
    }
    VX_MARK(5);
    hipLaunchKernelGGL(k_finalize, dim3((nb * nout + 255) / 256), dim3(256), 0, st, ws.keys, nb, nk, nout,
                       n, r0, best, trans);
    VX_LAUNCH_CHECK();
    VX_MARK(6);
#undef VX_MARK
    return VX_OK;
}

static int clear_if_dirty(vx_plan* p, hipStream_t st) {
    const size_t n3 = (size_t)p->n * p->n * p->n;
    if (p->q_dirty) {
        // the gather writes only the support box: clear the planes once
        VX_HIP(hipMemsetAsync(p->q, 0, (size_t)p->batch * p->C * n3 * sizeof(float), st));
        if (p->qd) VX_HIP(hipMemsetAsync(p->qd, 0, (size_t)p->batch * n3 * sizeof(double), st));
        if (p->q2) VX_HIP(hipMemsetAsync(p->q2, 0, (size_t)p->batch * p->C * n3 * sizeof(float), st));
        if (p->qd2) VX_HIP(hipMemsetAsync(p->qd2, 0, (size_t)p->batch * n3 * sizeof(double), st));
        p->q_dirty = false;
    }
    return VX_OK;
}

static int score_rows_impl(vx_plan* p, const float* rot, int nrows, float* best, int32_t* trans,
                           hipStream_t st, hipEvent_t* ev, double* stage_ms) {
    VX_REQUIRE(p && rot && best && trans, VX_ERR_ARG, "NULL argument");
    VX_REQUIRE(p->have_target && p->have_search, VX_ERR_STATE, "target/search not set");
    VX_REQUIRE(nrows >= 0, VX_ERR_ARG, "negative row count");
    int rc0 = clear_if_dirty(p, st);
    if (rc0 != VX_OK) return rc0;
    const Workspace ws0 = {p->q, p->qd, p->spec, p->sum, p->keys, p->real_out};
    const int nbatch = (nrows + p->batch - 1) / p->batch;
    if (ev || p->nstreams < 2 || nbatch < 2) {
        for (int r0 = 0; r0 < nrows; r0 += p->batch) {
            const int nb = std::min(p->batch, nrows - r0);
            int rc = score_batch(p, ws0, rot, r0, nb, best, trans, st, ev);
            if (rc != VX_OK) return rc;
            if (ev) {
                VX_HIP(hipEventSynchronize(ev[kStages]));
                for (int i = 0; i < kStages; ++i) {
                    float ms = 0.f;
                    VX_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                    stage_ms[i] += ms;
                }
            }
        }
        return VX_OK;
    }
    // Two batches in flight: even batches on the caller's stream with the
    // first workspace, odd ones on the plan's second stream with the second
    // (fork/join through events, so the caller's stream orders everything
    // and graph capture records the fork).
    const Workspace ws1 = {p->q2, p->qd2, p->spec2, p->sum2, p->keys2, p->real_out2};
    VX_HIP(hipEventRecord(p->ev_fork, st));
    VX_HIP(hipStreamWaitEvent(p->st2, p->ev_fork, 0));
    for (int i = 0; i < nbatch; ++i) {
        const int r0 = i * p->batch;
        const int nb = std::min(p->batch, nrows - r0);
        int rc = (i & 1) ? score_batch(p, ws1, rot, r0, nb, best, trans, p->st2, nullptr)
                         : score_batch(p, ws0, rot, r0, nb, best, trans, st, nullptr);
        if (rc != VX_OK) return rc;
    }
    VX_HIP(hipEventRecord(p->ev_join, p->st2));
    VX_HIP(hipStreamWaitEvent(st, p->ev_join, 0));
    return VX_OK;
}

extern "C" int vx_score_rows(vx_plan* p, const float* rot, int nrows, float* best, int32_t* trans,
                             vx_stream stream) {
    return score_rows_impl(p, rot, nrows, best, trans, (hipStream_t)stream, nullptr, nullptr);
}

// FP64 re-scoring workspace: grow-only, freed with the plan.
static int rs_grow(void** ptr, size_t bytes, size_t old_bytes, vx_plan* p) {
    if (*ptr && bytes <= old_bytes) return VX_OK;
    if (*ptr) { VX_HIP(hipFree(*ptr)); *ptr = nullptr; }
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        set_error("device allocation of %zu bytes for the FP64 re-scoring workspace failed", bytes);
        return VX_ERR_NOMEM;
    }
    p->bytes += bytes - (old_bytes < bytes ? old_bytes : 0);
    return VX_OK;
}

extern "C" int vx_rescore_f64(vx_plan* p, const double* tdata, const float* tvec, const float* tprob, double t_ave,
                              const double* sdata, const float* svec, const float* sprob, const float* rot,
                              int nrows, double window, int max_cand, double* score, int32_t* trans,
                              int32_t* ncand, vx_stream stream) {
    VX_REQUIRE(p && rot && score && trans && ncand, VX_ERR_ARG, "NULL argument");
    VX_REQUIRE(p->have_target && p->have_search, VX_ERR_STATE, "target/search not set");
    VX_REQUIRE(nrows >= 0, VX_ERR_ARG, "negative row count");
    VX_REQUIRE(window >= 0.0 && window < 1.0, VX_ERR_ARG, "window must be in [0, 1)");
    VX_REQUIRE(max_cand >= 1 && max_cand <= 4096, VX_ERR_ARG, "max_cand must be in [1, 4096]");
    VX_REQUIRE(sdata, VX_ERR_ARG, "sdata_dev is required (density masks)");
    const bool prob = p->mode == VX_MODE_PROB;
    if (p->mode == VX_MODE_VEC || prob) VX_REQUIRE(tvec && svec, VX_ERR_ARG, "tvec_dev and svec_dev required");
    if (prob) VX_REQUIRE(tprob && sprob, VX_ERR_ARG, "tprob_dev and sprob_dev required in prob mode");
    if (p->mode != VX_MODE_VEC && !prob) VX_REQUIRE(tdata, VX_ERR_ARG, "tdata_dev required in this mode");
    if (nrows == 0) return VX_OK;
    // the candidates come from the FFT x pass's float32 maps whatever x pass
    // the search runs (the MFMA forms round differently)
    struct XpassGuard {
        vx_plan* p;
        int saved;
        ~XpassGuard() { p->xvariant = saved; }
    } guard{p, p->xvariant};
    if (p->xvariant == 2 || p->xvariant == 4) p->xvariant = p->n * 4 <= 256 * kXcorrPF ? 3 : 1;
    hipStream_t st = (hipStream_t)stream;
    const int n = p->n;
    const size_t n3 = (size_t)n * n * n;
    const int gout = prob ? (p->mix ? 3 : 2) : 1;
    const size_t slots = (size_t)nrows * gout;
    int rc = rs_grow((void**)&p->rs_maps, (size_t)p->batch * p->G * n3 * sizeof(float),
                     p->rs_maps_cap * sizeof(float), p);
    if (rc != VX_OK) return rc;
    p->rs_maps_cap = std::max(p->rs_maps_cap, (size_t)p->batch * p->G * n3);
    if (nrows > p->rs_rows_alloc || max_cand > p->rs_mc_alloc) {
        const int rows = std::max(nrows, p->rs_rows_alloc);
        const int mc = std::max(max_cand, p->rs_mc_alloc);
        const size_t sl = (size_t)rows * 3;   // up to 3 groups
        const size_t osl = (size_t)p->rs_rows_alloc * 3;
        rc = rs_grow((void**)&p->rs_best, sl * sizeof(float), osl * sizeof(float), p);
        if (!rc) rc = rs_grow((void**)&p->rs_trans, sl * 3 * sizeof(int32_t), osl * 3 * sizeof(int32_t), p);
        if (!rc) rc = rs_grow((void**)&p->rs_cand, sl * mc * sizeof(int), osl * p->rs_mc_alloc * sizeof(int), p);
        if (!rc) rc = rs_grow((void**)&p->rs_partial, rs_partial_doubles(rows, 3, mc) * sizeof(double),
                              rs_partial_doubles(p->rs_rows_alloc, 3, p->rs_mc_alloc) * sizeof(double), p);
        if (rc != VX_OK) return rc;
        p->rs_rows_alloc = rows;
        p->rs_mc_alloc = mc;
    }
    rc = clear_if_dirty(p, st);
    if (rc != VX_OK) return rc;
    VX_HIP(hipMemsetAsync(ncand, 0, slots * sizeof(int32_t), st));
    const Workspace ws0 = {p->q, p->qd, p->spec, p->sum, p->keys, p->real_out};
    const RsStats stats = {p->alpha, p->vstd, p->vave, p->pstd, p->pave};
    const RsInputs in = {p->mode, tdata, tvec, tprob, t_ave, sdata, svec, sprob, rot};
    const int fma = (p->flags & VX_FLAG_NO_FMA) ? 0 : 1;
    for (int r0 = 0; r0 < nrows; r0 += p->batch) {
        const int nb = std::min(p->batch, nrows - r0);
        rc = score_batch(p, ws0, rot, r0, nb, p->rs_best, p->rs_trans, st, nullptr, p->rs_maps);
        if (rc != VX_OK) return rc;
        rc = rs_launch_candidates(p->rs_maps, p->G, n, nb, gout, r0, p->rs_best, window, stats, max_cand, ncand,
                                  p->rs_cand, st);
        if (rc != VX_OK) return rc;
        rc = rs_launch_direct(in, n, p->lo, p->w, fma, gout, max_cand, r0, nb, ncand, p->rs_cand, p->rs_partial,
                              st);
        if (rc != VX_OK) return rc;
    }
    return rs_launch_select(nrows, gout, n, max_cand, prob ? 1 : 0, ncand, p->rs_cand, p->rs_partial, stats,
                            score, trans, st);
}

// Per-channel correlation maps against caller-held conjugated spectra:
// out_c = irfftn(tspec_c * rfftn(q_c)) unnormalised, for each channel c.
extern "C" int vx_correlate(vx_plan* p, const void* tspec, int nch, const float* q, float* out,
                            vx_stream stream) {
    VX_REQUIRE(p && tspec && q && out && nch >= 1, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const int n = p->n;
    const size_t n3 = (size_t)n * n * n;
    const size_t nh = (size_t)n * n * p->nhz;
    const float2* ts = (const float2*)tspec;
    const int xv = p->xvariant == 0 ? 0 : 1;   // the FFT x passes valid for every grid edge
    for (int c = 0; c < nch; ++c) {
        const float* qc = q + (size_t)c * n3;
        int rc = p->pencil ? launch_fwd_pencil(qc, p->spec, p->fd, n, 1, 0, n, st, p->static_plans != 0, p->slab_fused != 0)
                           : launch_fwd_zy(qc, p->spec, p->fd, n, 1, 0, n, st, 0, p->static_plans != 0,
                                           p->slab_fused != 0);
        if (rc != VX_OK) return rc;
        rc = launch_xcorr(p->spec, ts + (size_t)c * nh, p->sum, p->fd, n, 1, 1, 1, 0u, 0, n, st, xv, p->xT,
                          p->static_plans != 0);
        if (rc != VX_OK) return rc;
        rc = p->pencil ? launch_inv_pencil(p->sum, p->fd, n, 1, 1, p->keys, 1, out + (size_t)c * n3, st, p->static_plans != 0,
                                           p->slab_fused != 0, p->pencil_chunks)
                       : launch_inv_yz(p->sum, p->fd, n, 1, 1, p->keys, 1, out + (size_t)c * n3, st, 0,
                                       p->static_plans != 0, p->slab_fused != 0);
        if (rc != VX_OK) return rc;
    }
    // the q planes were not touched, but the keys were: a later search re-zeroes them per batch
    return VX_OK;
}

extern "C" int vx_score_rows_profile(vx_plan* p, const float* rot, int nrows, float* best, int32_t* trans,
                                     double* stage_ms, vx_stream stream) {
    VX_REQUIRE(stage_ms, VX_ERR_ARG, "NULL stage_ms");
    hipEvent_t ev[kStages + 1];
    for (int i = 0; i <= kStages; ++i) VX_HIP(hipEventCreate(&ev[i]));
    for (int i = 0; i < kStages; ++i) stage_ms[i] = 0.0;
    const int rc = score_rows_impl(p, rot, nrows, best, trans, (hipStream_t)stream, ev, stage_ms);
    for (int i = 0; i <= kStages; ++i) (void)hipEventDestroy(ev[i]);
    return rc;
}

// ---------------------------------------------------------------------------
// Single-rotation map rotation (rot_mrc / rot_mrc_prob API)
// ---------------------------------------------------------------------------

namespace vx {

__global__ void __launch_bounds__(256)
k_rotate_map(int n, int nzb, RotF rf, int fma, const double* __restrict__ data, const float* __restrict__ vec,
             const float* __restrict__ prob, double* __restrict__ odata, float* __restrict__ ovec,
             float* __restrict__ oprob) {
    const int x = blockIdx.y;
    const int y = (blockIdx.x / nzb) * 4 + (threadIdx.x >> 6);
    const int z = (blockIdx.x % nzb) * 64 + (threadIdx.x & 63);
    if (y >= n || z >= n) return;
    const size_t n3 = (size_t)n * n * n;
    const float* F = rf.m;
    int cx[2], cy[2], cz[2];
    const int nx = dest_candidates(x, n, cx);
    const int ny = dest_candidates(y, n, cy);
    const int nz = dest_candidates(z, n, cz);
    const float cent = 0.5f * (float)n;
    const float fn = (float)n;
    double od = 0.0;
    float ov[3] = {0.f, 0.f, 0.f};
    float op[4] = {0.f, 0.f, 0.f, 0.f};
    bool done_v = false;
    bool done_p[4] = {prob == nullptr, prob == nullptr, prob == nullptr, prob == nullptr};
    for (int a = 0; a < nz; ++a)
    for (int b = 0; b < nx; ++b)
    for (int c = 0; c < ny; ++c) {
        const float n0 = (float)cx[b] - cent, n1 = (float)cy[c] - cent, n2 = (float)cz[a] - cent;
        const float o0 = rot_coord(n0, n1, n2, F + 0, cent, fma != 0);
        const float o1 = rot_coord(n0, n1, n2, F + 3, cent, fma != 0);
        const float o2 = rot_coord(n0, n1, n2, F + 6, cent, fma != 0);
        if (!(o0 >= 0.f && o1 >= 0.f && o2 >= 0.f && o0 < fn && o1 < fn && o2 < fn)) continue;
        const size_t si = ((size_t)(int)o0 * n + (int)o1) * n + (int)o2;
        if (!done_v && data[si] != 0.0) {
            od = data[si];
            const float s0 = vec[3 * si], s1 = vec[3 * si + 1], s2 = vec[3 * si + 2];
            ov[0] = rot_vec(s0, s1, s2, F, 0, fma != 0);
            ov[1] = rot_vec(s0, s1, s2, F, 1, fma != 0);
            ov[2] = rot_vec(s0, s1, s2, F, 2, fma != 0);
            done_v = true;
        }
        for (int k = 0; k < 4; ++k) {
            if (!done_p[k]) {
                const float pv = prob[k * n3 + si];
                if (pv != 0.f) { op[k] = pv; done_p[k] = true; }
            }
        }
        if (done_v && done_p[0] && done_p[1] && done_p[2] && done_p[3]) break;
    }
    const size_t v = ((size_t)x * n + y) * n + z;
    odata[v] = od;
    ovec[3 * v] = ov[0];
    ovec[3 * v + 1] = ov[1];
    ovec[3 * v + 2] = ov[2];
    if (oprob)
        for (int k = 0; k < 4; ++k) oprob[k * n3 + v] = op[k];
}

}  // namespace vx

extern "C" int vx_rotate_map(int n, const float* rot9, const double* data, const float* vec, const float* prob,
                             double* odata, float* ovec, float* oprob, int flags, vx_stream stream) {
    VX_REQUIRE(n > 0 && rot9 && data && vec && odata && ovec, VX_ERR_ARG, "NULL argument");
    VX_REQUIRE((prob == nullptr) == (oprob == nullptr), VX_ERR_ARG, "prob and out_prob must both be set or NULL");
    RotF rf;
    memcpy(rf.m, rot9, sizeof(rf.m));
    const int nzb = (n + 63) / 64;
    dim3 grid(nzb * ((n + 3) / 4), n);
    hipLaunchKernelGGL(k_rotate_map, grid, dim3(256), 0, (hipStream_t)stream, n, nzb, rf,
                       (flags & VX_FLAG_NO_FMA) ? 0 : 1, data, vec, prob, odata, ovec, oprob);
    VX_LAUNCH_CHECK();
    return VX_OK;
}

// The query planes vx_score_rows correlates for one rotation: the search's
// own gather (k_gather, plus the Laplacian for that mode) into out_dev, with
// zeros outside the plan's support box.
extern "C" int vx_query_planes(vx_plan* p, const float* rot_dev, float* out_dev, vx_stream stream) {
    VX_REQUIRE(p && rot_dev && out_dev, VX_ERR_ARG, "NULL argument");
    VX_REQUIRE(p->have_search, VX_ERR_STATE, "search map not set");
    hipStream_t st = (hipStream_t)stream;
    const int n = p->n;
    const size_t n3 = (size_t)n * n * n;
    VX_HIP(hipMemsetAsync(out_dev, 0, (size_t)p->C * n3 * sizeof(float), st));
    if (p->kind == KIND_DATA64) VX_HIP(hipMemsetAsync(p->qd, 0, n3 * sizeof(double), st));
    int rc = launch_gather(p->kind, p->svec, p->sscal, p->sprob, p->sdat, rot_dev, n, 1,
                           (p->flags & VX_FLAG_NO_FMA) ? 0 : 1, out_dev, p->qd, nullptr, 0, p->lo, p->w, p->src_lo,
                           p->src_w, st);
    if (rc != VX_OK) return rc;
    if (p->kind == KIND_DATA64) {
        hipLaunchKernelGGL(k_lap_to_q, dim3((unsigned)((n3 + 255) / 256), 1), dim3(256), 0, st, p->qd, out_dev, n);
        VX_LAUNCH_CHECK();
        // the search's own planes are cleared before its next batch
        p->q_dirty = true;
    }
    return VX_OK;
}
