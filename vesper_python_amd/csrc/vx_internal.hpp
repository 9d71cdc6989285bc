// Declarations shared by the translation units of libvesper_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <vector>

#include "vx_fft.hpp"

namespace vx {

// LDS available to one workgroup on gfx950 (160 KiB), minus room for the
// kernels' static shared arrays.
constexpr size_t kMaxLds = 160 * 1024 - 512;

struct RotF { float m[9]; };

// Raw buffer resource over `bytes` bytes from `base` (uniform across the
// workgroup); loads at offsets past the range return zero.  Third dword as
// the gfx9 family (gfx950) takes it.  Offsets are 32-bit, in bytes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float2 buf_load_f2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
    const u32x2_t v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

// Workgroup -> (column tile, rotation) of the x pass.  Workgroups are handed
// to the 8 XCDs round-robin by linear id, so XCD k = id % 8 is given the
// contiguous tile range [k tpx, (k + 1) tpx) and, within it, every rotation
// of a tile back to back: neighbouring tiles share 128-byte lines of the
// spectra and the target tile is reused by all rotations from that XCD's L2.
__device__ __forceinline__ bool xcd_tile(int id, int nb, int ntiles, int* tile, int* b) {
    const int tpx = (ntiles + 7) >> 3;
    const int j = id >> 3;
    const int tl = j / nb;
    *b = j - tl * nb;
    *tile = (id & 7) * tpx + tl;
    return tl < tpx && *tile < ntiles;
}

// Gather kinds (what a destination voxel collects from the source map).
enum { KIND_VEC = 0, KIND_SCALAR = 1, KIND_PROB = 2, KIND_DATA64 = 3 };

int fft_desc_create(int n, FftDesc* d, float2** twid_out, int order = 0);
int launch_gather(int kind, const float4* svec, const float2* sscal, const float4* sprob,
                  const double* sdat, const float* rot, int n, int nb, int fma, float* q,
                  double* qd, unsigned long long* keys, int nkeys, int lo, int w, int slo, int sw,
                  hipStream_t st);

// Fused slab kernels on compile-time plans (vx_slab.hip): each returns false
// (nothing launched) where the grid edge has no compile-time plan in its
// slab shape, else true with *rc the launch status.
bool launch_fwd_fused(const float* q, float2* spec, const FftDesc& fd, int n, int nplanes, int lo, int w,
                      hipStream_t st, int nt_pref, int* rc);
bool launch_inv_fused(const float2* sum, const FftDesc& fd, int n, int nbg, int G, unsigned long long* keys,
                      int nkeys, float* real_out, hipStream_t st, int nt_pref, int* rc);
// The pencil passes' fused forms (grid edges 201..400); launch_axis_y_fused
// covers the full column tiles (*full_tiles of them) and leaves the rest.
bool launch_fwd_z_fused(const float* q, float2* spec, const FftDesc& fd, int n, int nplanes, int lo, int w,
                        hipStream_t st, int* rc);
bool launch_axis_y_fused(float2* buf, const FftDesc& fd, int n, int inverse, int ylo, int yw, int xlo, int nxs,
                         int nplanes, hipStream_t st, int* full_tiles, int* rc);
// chunks: line-pair chunks per workgroup of the inverse z pass (0 = automatic)
bool launch_inv_z_fused(const float2* sum, const FftDesc& fd, int n, int nbg, int G, unsigned long long* keys,
                        int nkeys, float* real_out, hipStream_t st, int* rc, int chunks);

// Top-N selection (vx_score.hip): keys per ping-pong buffer, argument check,
// and the selection on caller-provided buffers of topn_cap keys each.
size_t topn_cap(int nrows, int topn);
int topn_check(const float* scores, int nrows, int stride, int col, int topn, const int32_t* rows);
int topn_run(const float* scores, int nrows, int stride, int col, int topn, int32_t* rows,
             unsigned long long* bufA, unsigned long long* bufB, hipStream_t st);

// FP64 re-scoring (vx_rescore.hip).
struct RsStats { double alpha, vstd, vave, pstd, pave; };
struct RsInputs {
    int mode;
    const double* tdata;
    const float* tvec;
    const float* tprob;
    double ave;
    const double* sdata;
    const float* svec;
    const float* sprob;
    const float* rot;
};
int rs_launch_candidates(const float* maps, int G, int n, int nb, int gout, int r0, const float* best32,
                         double window, const RsStats& st, int max_cand, int* ncand, int* cand, hipStream_t s);
int rs_launch_direct(const RsInputs& in, int n, int lo, int w, int fma, int gout, int max_cand, int r0, int nb,
                     const int* ncand, const int* cand, double* partial, hipStream_t s);
int rs_launch_select(int nrows, int gout, int n, int max_cand, int prob, const int* ncand, const int* cand,
                     const double* partial, const RsStats& st, double* score, int32_t* trans, hipStream_t s);
size_t rs_partial_doubles(int nrows, int gout, int max_cand);

// x pass with the length-N1 DFTs of N = N1 N2 on split-f16 MFMA (vx_xmr.hip);
// wtab holds W_N^j, j < N.
bool xcorr_mr_factor(int n, int* n1, int* n2);
int xcorr_mr_tables(int n, int lo, int w, std::vector<uint32_t>& out);
int launch_xcorr_mr(const float2* spec, const float2* tgt, float2* out, const float2* wtab, const void* atab, int n,
                    int nb, int C, int lo, int w, hipStream_t st);

// Ordered compaction of the rows whose prediction falls in histogram bins >=
// bin (vx_propose.hip); scratch holds 257 unsigned.
int compact_top_bins(int n, const float* v, const unsigned* mm, int bin, int32_t* out, unsigned* scratch,
                     hipStream_t st);

}  // namespace vx
