// tune.hip -- launch-parameter autotuner: times every candidate setting of a
// context on a sample batch and keeps the fastest one whose result bitmaps
// equal those of the settings the context held on entry.
//
// Knobs (all on the device-dispatched batch path):
//   dense group width  queries that share one pass over the value-tiled key
//                      stream (1..256); sets the LDS row-tile width (group x
//                      tile bits) and the query-mask word (u32, u64 or two
//                      u64 for more than 64 queries)
//   dense threads      work-group size of the persistent kernel (256 / 512),
//                      i.e. waves per value tile
//   sparse chunk       records per sparse route chunk (4096 / 8192; larger
//                      chunks give longer per-slice runs of references)
//   dense columns      per-value query masks in LDS (on / off)
// The value tile itself is a layout parameter of the index (ab_index_build's
// tile_shift); tuning it means building the index once per candidate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "engine.h"

using namespace ab;

namespace {

// Words that differ between two bitmap sets of n4 uint4 each; one atomic per wave.
__global__ void __launch_bounds__(256) k_count_diff(const uint4 *__restrict__ a, const uint4 *__restrict__ b,
                                                    uint64_t n4, unsigned long long *__restrict__ out)
{
    uint32_t d = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 x = a[i], y = b[i];
        d += (x.x != y.x) + (x.y != y.y) + (x.z != y.z) + (x.w != y.w);
    }
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
    if ((threadIdx.x & 63) == 0 && d) atomicAdd(out, (unsigned long long)d);
}

struct TuneScratch {
    void *ref = nullptr, *cand = nullptr, *diff = nullptr;
    ~TuneScratch()
    {
        if (ref) (void)hipFree(ref);
        if (cand) (void)hipFree(cand);
        if (diff) (void)hipFree(diff);
    }
};

}  // namespace

extern "C" ab_status ab_ctx_autotune(ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq,
                                     int path, uint32_t repeats, ab_tune_config *out, uint32_t cap,
                                     uint32_t *n_out)
{
    if (!c || !h_qoff || nq == 0 || (h_qoff[nq] && !d_keys) || repeats == 0 || (cap && !out)) {
        set_error("ab_ctx_autotune: bad argument");
        return AB_INVALID_ARG;
    }
    const uint32_t groups[] = {8, 16, 32, 48, 64, 96, 128, 192, 256};
    const uint32_t threads[] = {256, 512};
    const uint32_t chunks[] = {4096, 8192};
    const int columns[] = {1, 0};
    std::vector<ab_tune_config> all;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("ab_ctx_autotune: hipEventCreate failed");
        return AB_DEVICE_ERROR;
    }
    const uint32_t saved[4] = {c->group, c->dense_threads, c->sparse_chunk, (uint32_t)c->dense_columns};
    // the candidates' settings must hold for every call: the batch-size table
    // is set aside until the end
    struct KeepTable {
        ab_ctx *c;
        std::vector<ab_tuning> t;
        explicit KeepTable(ab_ctx *cc) : c(cc) { t.swap(c->tuning); }
        ~KeepTable() { c->tuning.swap(t); }
    } keep(c);
    // result bitmaps of the entry settings: every candidate must reproduce them
    const uint64_t bm_bytes = (uint64_t)nq * c->ix->stride;  // stride: a multiple of 16 bytes
    TuneScratch sc;
    if (hipMalloc(&sc.ref, bm_bytes) != hipSuccess || hipMalloc(&sc.cand, bm_bytes) != hipSuccess ||
        hipMalloc(&sc.diff, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipGetLastError();
        set_error("ab_ctx_autotune: out of device memory for the result check (" + std::to_string(2 * bm_bytes) +
                  " bytes)");
        return AB_OOM;
    }
    ab_status st = ab_query_batch(c, d_keys, h_qoff, nq, path, (uint32_t *)sc.ref, nullptr, 0, nullptr, nullptr);
    const uint32_t diff_grid = (uint32_t)std::min<uint64_t>(4096, (bm_bytes / 16 + 255) / 256);
    for (uint32_t g : groups)
        for (uint32_t t : threads)
            for (uint32_t ch : chunks)
                for (int col : columns) {
                    if (st != AB_OK) break;
                    if ((st = ab_ctx_set_dense_group(c, g)) != AB_OK) break;
                    if ((st = ab_ctx_set_dense_threads(c, t)) != AB_OK) break;
                    if ((st = ab_ctx_set_sparse_chunk(c, ch)) != AB_OK) break;
                    if ((st = ab_ctx_set_dense_columns(c, col)) != AB_OK) break;
                    // checked warm-up (allocation, graph capture), then timed repeats
                    st = ab_query_batch(c, d_keys, h_qoff, nq, path, (uint32_t *)sc.cand, nullptr, 0, nullptr,
                                        nullptr);
                    if (st != AB_OK) break;
                    unsigned long long diff = 0;
                    if (hipMemsetAsync(sc.diff, 0, sizeof(diff), c->stream) != hipSuccess) st = AB_DEVICE_ERROR;
                    if (st == AB_OK) {
                        k_count_diff<<<diff_grid, 256, 0, c->stream>>>((const uint4 *)sc.ref, (const uint4 *)sc.cand,
                                                                       bm_bytes / 16, (unsigned long long *)sc.diff);
                        if (hipGetLastError() != hipSuccess ||
                            hipMemcpyAsync(&diff, sc.diff, sizeof(diff), hipMemcpyDeviceToHost, c->stream) !=
                                hipSuccess ||
                            hipStreamSynchronize(c->stream) != hipSuccess)
                            st = AB_DEVICE_ERROR;
                    }
                    if (st != AB_OK) {
                        set_error("ab_ctx_autotune: result check failed");
                        break;
                    }
                    (void)hipEventRecord(e0, c->stream);
                    for (uint32_t r = 0; r < repeats && st == AB_OK; ++r)
                        st = ab_query_batch(c, d_keys, h_qoff, nq, path, nullptr, nullptr, 0, nullptr, nullptr);
                    (void)hipEventRecord(e1, c->stream);
                    float ms = 0;
                    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                        set_error("ab_ctx_autotune: event timing failed");
                        st = AB_DEVICE_ERROR;
                    }
                    ab_tune_config cfg{};
                    cfg.group = max_group(c);  // the width actually used after the LDS limit
                    cfg.dense_threads = t;
                    cfg.sparse_chunk = ch;
                    cfg.dense_columns = col && col_fits(c, group_mode(cfg.group), cfg.group);
                    cfg.ms = (double)ms / repeats;
                    cfg.mismatched_words = diff;
                    all.push_back(cfg);
                }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != AB_OK || all.empty()) {
        (void)ab_ctx_set_dense_group(c, saved[0]);
        (void)ab_ctx_set_dense_threads(c, saved[1]);
        (void)ab_ctx_set_sparse_chunk(c, saved[2]);
        (void)ab_ctx_set_dense_columns(c, (int)saved[3]);
        return st != AB_OK ? st : AB_DEVICE_ERROR;
    }
    // valid candidates first, each part fastest first
    std::stable_sort(all.begin(), all.end(), [](const ab_tune_config &a, const ab_tune_config &b) {
        if ((a.mismatched_words == 0) != (b.mismatched_words == 0)) return a.mismatched_words == 0;
        return a.ms < b.ms;
    });
    if (all.front().mismatched_words) {  // no candidate reproduced the entry settings' bitmaps
        (void)ab_ctx_set_dense_group(c, saved[0]);
        (void)ab_ctx_set_dense_threads(c, saved[1]);
        (void)ab_ctx_set_sparse_chunk(c, saved[2]);
        (void)ab_ctx_set_dense_columns(c, (int)saved[3]);
        set_error("ab_ctx_autotune: every candidate's bitmaps differ from the entry settings' result");
        return AB_DEVICE_ERROR;
    }
    const ab_tune_config &best = all.front();
    (void)ab_ctx_set_dense_group(c, best.group);
    (void)ab_ctx_set_dense_threads(c, best.dense_threads);
    (void)ab_ctx_set_sparse_chunk(c, best.sparse_chunk);
    (void)ab_ctx_set_dense_columns(c, best.dense_columns);
    const uint32_t n = std::min<uint32_t>(cap, (uint32_t)all.size());
    for (uint32_t i = 0; i < n; ++i) out[i] = all[i];
    if (n_out) *n_out = (uint32_t)all.size();
    return AB_OK;
}
