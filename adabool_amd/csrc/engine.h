// engine.h -- host-side engine objects shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.h"
#include "kernels.h"

namespace ab {

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    ab_status ensure(size_t bytes);
    void release();
};

// Grow-only pinned host buffer.
struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    ab_status ensure(size_t bytes);
    void release();
};

// Arguments of one batched evaluation.
struct BatchArgs {
    const uint32_t *d_keys = nullptr;  // concatenated keys (device)
    const uint64_t *h_qoff = nullptr;  // nq+1 offsets (host)
    uint32_t nq = 0;
    int path = AB_PATH_THRESHOLD;
    const uint8_t *h_sel = nullptr;   // per-query path decided on the host, or NULL
    const uint64_t *h_M = nullptr;    // per-query matched records with h_sel, or NULL
    uint32_t *d_bitmaps = nullptr;    // NULL: context scratch
    uint64_t out_stride = 0;          // u32 words between consecutive queries' bitmaps
    uint64_t out_words = 0;           // words written per query
    uint32_t *d_ids = nullptr;
    uint64_t ids_stride = 0;
    const uint64_t *h_id_off = nullptr;  // per-query id offsets (host), replacing ids_stride when set
    bool id_off_aligned = false;         // every offset is a multiple of 4
    uint32_t *d_counts = nullptr;
    uint8_t *h_paths = nullptr;
};

// MLP cost model: 12 -> 64 -> 64 -> 2, FP32, standardized inputs.
constexpr int kF = AB_NFEATURES, kH1 = 64, kH2 = 64, kO = 2;
constexpr int kParams = kH1 * kF + kH1 + kH2 * kH1 + kH2 + kO * kH2 + kO;

}  // namespace ab

struct ab_index {
    uint32_t nkeys = 0, nvalues = 0;
    uint64_t nnz = 0;
    uint64_t nrec[AB_NTYPES] = {0};
    uint64_t R = 0, list_bytes = 0;
    uint32_t tile_shift = 0, ntiles = 0;
    uint64_t nblocks = 0, kbytes = 0, stride = 0;
    std::vector<uint32_t> seg8;  // host copy for host-side dispatch
    std::vector<void *> allocs;
    uint64_t dev_bytes = 0;
    ab::DevIndex dev{};
    int device = 0;
    bool host_only = false;
    ab::HostIndex host;  // full host layout (host-only handles)
};

struct ab_ctx {
    ab_index *ix = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double tau = 0.05;          // threshold rule for single queries
    double tau_batch = 0.001;   // threshold rule inside batches (dense passes are shared)
    uint32_t group = 64;
    uint32_t dense_threads = 512;
    bool dense_columns = true;  // per-value query masks in dense groups when LDS allows
    uint32_t sparse_chunk = 4096;
    // device scratch
    ab::DevBuf qoff, sel, kq, cnt, koff, parts, feat, M, ccnt, bm, qlist, keys, qbm, mask, idoff;
    // pinned staging: control data (qoff, decisions, query lists) and keys
    ab::PinBuf hstage, hkstage, hsel, hM;
    hipEvent_t stage_ev = nullptr, kstage_ev = nullptr;
    bool stage_pending = false, kstage_pending = false;
    // device-dispatched batches (ab_query_batch):
    //   0 host-dispatched launches, 1 one fused persistent union launch with
    //   device-side dispatch and grouping, 2 = 1 replayed from a cached hipGraph
    int batch_mode = 2;
    uint64_t cfg_gen = 1;  // bumped by every setting that changes launch parameters
    ab::DevBuf qgroup, grp_q, grp_n, plan, counter, gmask, ganybm;
    // launch settings by batch size (ab_ctx_set_batch_tuning), ascending min_batch
    std::vector<ab_tuning> tuning;
    // id-only batches (ab_query_batch_ids): values expanded per part of at
    // most id_budget values, chunk counts and bases, sort output, run starts,
    // rocPRIM temporary storage
    uint64_t id_budget = 1ull << 27;
    int id_method = 0;  // 0: by value slice where it applies, 1: always by sort
    ab::DevBuf id_vals, id_sorted, id_ccnt, id_cbase, id_start, id_tmp;
    // id-only batches by value slice: per-part route arrays (device and their
    // pinned staging), the work-groups of the slice kernel
    ab::DevBuf idpart;
    ab::PinBuf hidpart;
    uint32_t ids_grid = 0;
    uint32_t gmask_groups = 0;  // groups the zeroed mask buffers hold
    int gmask_mode = -1;
    struct GraphCache {
        bool valid = false;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<uint64_t> key;
        std::vector<uint64_t> pending;  // the last shape run without a graph
    } gc;
    ab::DevBuf gkeys;  // graph mode: the batch's keys, copied in before every launch
    // write-once sparse work of device-dispatched batches: per-query flags and
    // prefixes, routed record refs (at most sp_cap) and per-chunk slice offsets
    ab::DevBuf sp_flag, sp_cpre, sp_scpre, sp_rbase, sp_slist, sp_tot, sp_refs, sp_choff;
    uint64_t sp_cap = 1ull << 28;
    uint32_t sp_grid = 0;  // work-groups of the route launch
    uint32_t fused_grid = 0;  // work-groups of the last fused launch
    // per-kernel-class timing
    bool timing = false;
    struct Span {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Span> spans;
    std::vector<hipEvent_t> ev_pool;
    double t_ms[AB_K_NCLASSES] = {0};
    uint64_t t_n[AB_K_NCLASSES] = {0};
    // cost models: slot 0 prices single queries (wall time of one call),
    // slot 1 prices a query's share of a batch (device time of its kernels)
    struct Model {
        bool loaded = false;
        std::vector<float> host;  // params followed by mean[12], std[12]
        ab::DevBuf params, norm;
    };
    Model mlp[2];
    ~ab_ctx();
};

namespace ab {
// Records a start event at construction and an end event at destruction on
// the context stream when timing is on.
struct TimedSpan {
    ab_ctx *c;
    int kind;
    hipEvent_t a = nullptr;
    TimedSpan(ab_ctx *ctx, int k);
    ~TimedSpan();
};
ab_status run_batch(ab_ctx *c, const BatchArgs &a);
// Device-dispatched batch: one fused persistent union launch, optionally
// replayed from a cached hipGraph (fused.hip).
ab_status run_batch_fused(ab_ctx *c, const BatchArgs &a);
void release_graph(ab_ctx *c);
// Upper bound of a query's id count: the sum of its in-range keys' list
// lengths, capped at nvalues (host seg8 copy).
uint64_t id_upper_bound(const ab_index *ix, const uint32_t *keys, uint64_t n);
// Largest dense group width that fits the LDS budget and whether its
// per-value column masks fit as well.
uint32_t max_group(const ab_ctx *c);
int group_mode(uint32_t B);
bool col_fits(const ab_ctx *c, int mode, uint32_t B);
ab_status stage_wait(ab_ctx *c);
// Device decision for nq queries from device features with model `slot`
// (MLP forward kernel).
ab_status mlp_decide(ab_ctx *c, int slot, const float *d_feat, uint32_t nq, uint8_t *d_sel);
// Host decision for one query's features with model `slot`.
int mlp_decide_host(const ab_ctx *c, int slot, const float *f);
// Looks up the no-allocation entry point's per-device scratch address
// (noalloc.hip); done once per device, at context creation.
ab_status noalloc_setup(int dev);
}  // namespace ab
