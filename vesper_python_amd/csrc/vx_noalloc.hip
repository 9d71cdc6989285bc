// vx_topn_noalloc: the top-N selection of vx_topn for callers that cannot
// allocate device memory at all (code under HIP-graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  Its key buffers live in a
// file-scope __device__ array that every GPU holds from the moment the code
// object loads until the process ends; nothing is allocated or freed.
#include <hip/hip_runtime.h>
#include <mutex>

#include "vx_common.hpp"
#include "vx_internal.hpp"

namespace vx {

// Largest row count a call may select from: 2^18 = 262,144 rows covers the
// 191,808 Euler rows of a 5-degree search.  Two ping-pong buffers of
// topn_cap(kNoAllocMaxRows, 1024) = 128 * 1024 + 2048 keys each: 2.1 MB.
constexpr int kNoAllocMaxRows = 1 << 18;
constexpr int kNoAllocSortN = 2048;     // rows per sorting chunk (kSortN of vx_score.hip)
constexpr size_t kNoAllocCap = (size_t)(kNoAllocMaxRows / kNoAllocSortN) * 1024 + kNoAllocSortN;

static __device__ __align__(256) unsigned long long scratch[2 * kNoAllocCap];

// Device address of `scratch` on each GPU.  In host code `scratch` names a
// host-side stand-in, not device memory: only these addresses may be handed
// to kernels.
constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static unsigned long long* g_noalloc_ptr[kNoAllocMaxDevices] = {};

static int noalloc_scratch(unsigned long long** out) {
    int dev = -1;
    VX_HIP(hipGetDevice(&dev));
    VX_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, VX_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        VX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VX_REQUIRE(p != nullptr, VX_ERR_HIP, "no device address for the top-N scratch on device %d", dev);
        g_noalloc_ptr[dev] = (unsigned long long*)p;
    }
    *out = g_noalloc_ptr[dev];
    return VX_OK;
}

}  // namespace vx

extern "C" int vx_topn_noalloc(const float* scores, int nrows, int stride, int col, int topn, int32_t* rows,
                               vx_stream stream) {
    int rc = vx::topn_check(scores, nrows, stride, col, topn, rows);
    if (rc) return rc;
    VX_REQUIRE(nrows <= vx::kNoAllocMaxRows, VX_ERR_ARG,
               "vx_topn_noalloc takes at most %d rows (got %d); use vx_topn or vx_topn_ws", vx::kNoAllocMaxRows,
               nrows);
    const size_t cap = vx::topn_cap(nrows, topn);
    if (cap > vx::kNoAllocCap) {
        vx::set_error("internal: top-N needs %zu keys per buffer, scratch holds %zu", cap, vx::kNoAllocCap);
        return VX_ERR_STATE;
    }
    unsigned long long* base = nullptr;
    rc = vx::noalloc_scratch(&base);
    if (rc) return rc;
    // every key a pass reads was written by the pass before it in this call,
    // so the buffers need no clearing between calls
    return vx::topn_run(scores, nrows, stride, col, topn, rows, base, base + vx::kNoAllocCap, (hipStream_t)stream);
}
