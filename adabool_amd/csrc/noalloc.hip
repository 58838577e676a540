// noalloc.hip -- the sparse method on device keys for callers that cannot
// allocate (code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc).  The plan, the scan and the union take their scratch
// from one fixed file-scope device array instead of context buffers; each GPU
// holds its own copy of it until the process ends.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.h"
#include "kernels.h"

namespace ab {

namespace {

// Keys per call; the scratch holds qoff[2], then kq (u32), cnt (u64),
// koff (u64, one more) and the scan's partial sums for that many keys.
constexpr uint64_t kNoAllocMaxKeys = 1ull << 18;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t kOffQoff = 0;
constexpr uint64_t kOffKq = align256(kOffQoff + 16);
constexpr uint64_t kOffCnt = align256(kOffKq + kNoAllocMaxKeys * 4);
constexpr uint64_t kOffKoff = align256(kOffCnt + kNoAllocMaxKeys * 8);
constexpr uint64_t kOffParts = align256(kOffKoff + (kNoAllocMaxKeys + 1) * 8);
// the scan needs ceil(n / 4096) + 1 partial sums (scan_parts_needed), plus
// two words of slack as the context path gives it
constexpr uint64_t kPartsBytes = ((kNoAllocMaxKeys + 4095) / 4096 + 1) * 8 + 16;
constexpr uint64_t kScratchBytes = align256(kOffParts + kPartsBytes);  // 5.25 MiB

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_looked[kMaxDevices] = {};

__global__ void k_noalloc_qoff(uint64_t *qoff, uint64_t n)
{
    if (threadIdx.x == 0) {
        qoff[0] = 0;
        qoff[1] = n;
    }
}

}  // namespace

// Looks up the device address of `scratch` once per device, with that device
// current.  In host code the array's name is only a host-side stand-in:
// kernels, copies and fills take this address instead.
ab_status noalloc_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("no-allocation scratch: device ordinal out of range");
        return AB_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_looked[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        g_err[dev] = e;
        g_ptr[dev] = e == hipSuccess ? p : nullptr;
        g_looked[dev] = true;
    }
    if (g_err[dev] != hipSuccess) {
        set_error(std::string("no-allocation scratch: hipGetSymbolAddress: ") + hipGetErrorString(g_err[dev]));
        return AB_DEVICE_ERROR;
    }
    return AB_OK;
}

static void *noalloc_ptr(int dev)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return (dev >= 0 && dev < kMaxDevices && g_looked[dev]) ? g_ptr[dev] : nullptr;
}

}  // namespace ab

using namespace ab;

extern "C" ab_status ab_list_union_sparse_noalloc(ab_ctx *c, const uint32_t *d_keys, uint32_t n, uint32_t *d_res,
                                                  uint64_t res_len_bytes)
{
    if (!c || (n && !d_keys) || !d_res) {
        set_error("ab_list_union_sparse_noalloc: bad argument");
        return AB_INVALID_ARG;
    }
    if (n > kNoAllocMaxKeys) {
        set_error("ab_list_union_sparse_noalloc: " + std::to_string(n) + " keys exceed the maximum of " +
                  std::to_string(kNoAllocMaxKeys));
        return AB_INVALID_ARG;
    }
    const ab_index *ix = c->ix;
    if (res_len_bytes < ix->stride) {
        set_error("ab_list_union_sparse_noalloc: res_len_bytes smaller than ab_result_stride_bytes()");
        return AB_INVALID_ARG;
    }
    // looked up when the context was created; a context made before a
    // failed lookup reports it here
    uint8_t *base = (uint8_t *)noalloc_ptr(ix->device);
    if (!base) {
        const ab_status st = noalloc_setup(ix->device);
        if (st != AB_OK) return st;
        base = (uint8_t *)noalloc_ptr(ix->device);
    }
    uint64_t *qoff = (uint64_t *)(base + kOffQoff);
    uint32_t *kq = (uint32_t *)(base + kOffKq);
    uint64_t *cnt = (uint64_t *)(base + kOffCnt);
    uint64_t *koff = (uint64_t *)(base + kOffKoff);
    uint64_t *parts = (uint64_t *)(base + kOffParts);
    hipStream_t s = c->stream;
    const uint64_t words = ix->stride / 4;
    // the result is written in full: zeros, then the union's OR-s; every
    // scratch word the kernels read is written earlier in the same call
    launch_fill_zero(d_res, ix->stride, s);
    hipLaunchKernelGGL(k_noalloc_qoff, dim3(1), dim3(64), 0, s, qoff, (uint64_t)n);
    if (n) {
        launch_sparse_plan(ix->dev, d_keys, qoff, nullptr, 1, kq, cnt, s);
        launch_scan_u64(cnt, koff, n, parts, s);
        launch_sparse_union(ix->dev, d_keys, kq, koff, n, d_res, words, c->sparse_chunk, 2048, s);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("ab_list_union_sparse_noalloc: ") + hipGetErrorString(e));
        return AB_DEVICE_ERROR;
    }
    return AB_OK;
}
