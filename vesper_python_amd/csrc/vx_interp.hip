// Parity-lattice interpolation of per-voxel class predictions
// (interpolate, interpolation_functions.py:6-135).
//
// The prediction grid [sx][sy][sz][nl] holds values on one parity lattice:
// the reference parity (rx, ry, rz) is that of the first voxel, in C order,
// whose channel 0 is positive.  Seven passes then fill the zero entries of
// the other lattices from their +-1 neighbours along the axes whose parity
// differs: the three single-axis lattices (x, y, z), then the three two-axis
// ones (xy, xz, yz), then the all-axis one.  Each pass reads only lattices
// written by earlier passes, so a pass is one parallel launch.  As in the
// reference:
//   - a pass covers indices start, start + 2, ... < dim - 1 on each axis;
//   - index -1 wraps to dim - 1 (Python/Numba negative indexing);
//   - the average is ((0 + a) + b ...) / m in float64, in the order
//     +x, -x, +y, -y, +z, -z over the differing axes, then stored in the
//     grid's own precision.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "vesper_hip.h"
#include "vx_common.hpp"

namespace vx {

template <typename T>
__global__ void __launch_bounds__(256)
k_first_positive(const T* __restrict__ d, size_t nvox, int nl, unsigned long long* __restrict__ first) {
    unsigned long long best = ~0ull;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x)
        if (d[v * nl] > (T)0) { best = v; break; }   // v only grows along a thread's stride
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off, 64);
        best = o < best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(first, best);
}

// One pass: mask bit a set = axis a's parity differs from the reference.
// Grid: (lattice z x channels in 256-thread blocks, lattice y, lattice x).
template <typename T>
__global__ void __launch_bounds__(256)
k_interp_pass(T* __restrict__ d, int sx, int sy, int sz, int nl, int x0, int y0, int z0, int cz, int mask) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= cz * nl) return;
    const int ck = t / nl, l = t - ck * nl;
    const int i = x0 + 2 * (int)blockIdx.z, j = y0 + 2 * (int)blockIdx.y, k = z0 + 2 * ck;
    const size_t sxs = (size_t)sy * sz * nl, sys = (size_t)sz * nl, szs = (size_t)nl;
    const size_t at = i * sxs + j * sys + k * szs + l;
    if (d[at] != (T)0) return;
    double avg = 0.0;
    int m = 0;
    if (mask & 1) {
        avg += (double)d[at + sxs];
        avg += (double)d[(i == 0) ? at + (size_t)(sx - 1) * sxs : at - sxs];
        m += 2;
    }
    if (mask & 2) {
        avg += (double)d[at + sys];
        avg += (double)d[(j == 0) ? at + (size_t)(sy - 1) * sys : at - sys];
        m += 2;
    }
    if (mask & 4) {
        avg += (double)d[at + szs];
        avg += (double)d[(k == 0) ? at + (size_t)(sz - 1) * szs : at - szs];
        m += 2;
    }
    d[at] = (T)(avg / (double)m);
}

template <typename T>
static int interpolate_t(T* d, int sx, int sy, int sz, int nl, hipStream_t st, unsigned long long* first_dev,
                         int* ref_out) {
    const size_t nvox = (size_t)sx * sy * sz;
    VX_HIP(hipMemsetAsync(first_dev, 0xff, sizeof(unsigned long long), st));
    const unsigned g0 = (unsigned)std::min<size_t>((nvox + 255) / 256, 4096);
    hipLaunchKernelGGL(k_first_positive<T>, dim3(g0), dim3(256), 0, st, d, nvox, nl, first_dev);
    VX_LAUNCH_CHECK();
    unsigned long long first = 0;
    VX_HIP(hipMemcpyAsync(&first, first_dev, sizeof(first), hipMemcpyDeviceToHost, st));
    VX_HIP(hipStreamSynchronize(st));
    VX_REQUIRE(first != ~0ull, VX_ERR_ARG, "no voxel with a positive channel 0: the reference parity is undefined");
    const int rz = (int)(first % sz) & 1, ry = (int)((first / sz) % sy) & 1, rx = (int)(first / ((size_t)sy * sz)) & 1;
    if (ref_out) { ref_out[0] = rx; ref_out[1] = ry; ref_out[2] = rz; }
    // pass order of the reference: x, y, z, xy, xz, yz, xyz
    const int masks[7] = {1, 2, 4, 3, 5, 6, 7};
    const int ref[3] = {rx, ry, rz};
    const int dims[3] = {sx, sy, sz};
    for (int pidx = 0; pidx < 7; ++pidx) {
        const int mask = masks[pidx];
        int start[3], cnt[3];
        for (int a = 0; a < 3; ++a) {
            start[a] = (mask >> a & 1) ? (ref[a] + 1) % 2 : ref[a];
            const int span = dims[a] - 1 - start[a];
            cnt[a] = span > 0 ? (span + 1) / 2 : 0;
        }
        if ((size_t)cnt[0] * cnt[1] * cnt[2] == 0) continue;
        VX_REQUIRE(cnt[0] <= 65535 && cnt[1] <= 65535 && (size_t)cnt[2] * nl < (1u << 30), VX_ERR_UNSUPPORTED,
                   "prediction grid too large");
        dim3 grid((unsigned)((cnt[2] * nl + 255) / 256), (unsigned)cnt[1], (unsigned)cnt[0]);
        hipLaunchKernelGGL(k_interp_pass<T>, grid, dim3(256), 0, st, d, sx, sy, sz, nl, start[0], start[1],
                           start[2], cnt[2], mask);
        VX_LAUNCH_CHECK();
    }
    return VX_OK;
}

}  // namespace vx

extern "C" int vx_interpolate(void* data, int is_f64, int sx, int sy, int sz, int nl, int32_t* ref_host3,
                              vx_stream stream) {
    VX_REQUIRE(data && sx > 0 && sy > 0 && sz > 0 && nl > 0, VX_ERR_ARG, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* first = nullptr;
    VX_HIP(hipMallocAsync((void**)&first, sizeof(unsigned long long), st));
    int ref[3] = {0, 0, 0};
    const int rc = is_f64 ? vx::interpolate_t((double*)data, sx, sy, sz, nl, st, first, ref)
                          : vx::interpolate_t((float*)data, sx, sy, sz, nl, st, first, ref);
    (void)hipFreeAsync(first, st);
    if (rc == VX_OK && ref_host3)
        for (int a = 0; a < 3; ++a) ref_host3[a] = ref[a];
    return rc;
}
