// fused.hip -- device-dispatched batches: features, dispatch decisions, dense
// grouping and query masks are computed on the device, then ONE persistent
// union launch takes dense (group, tile) items and sparse chunks from a
// shared work counter; ids are compacted after it.  Nothing returns to the
// host, so the whole sequence is captured once per batch shape and buffer set
// into a hipGraph and replayed (mode 2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace ab {

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));            \
            return e_ == hipErrorOutOfMemory ? AB_OOM : AB_DEVICE_ERROR;              \
        }                                                                             \
    } while (0)
#define AB_TRY(expr)                \
    do {                            \
        ab_status s_ = (expr);      \
        if (s_ != AB_OK) return s_; \
    } while (0)

__global__ void k_fill_u8(uint8_t *p, uint32_t n, uint8_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void release_graph(ab_ctx *c)
{
    if (c->gc.exec) (void)hipGraphExecDestroy(c->gc.exec);
    if (c->gc.graph) (void)hipGraphDestroy(c->gc.graph);
    c->gc = ab_ctx::GraphCache();
}

namespace {

struct Plan {
    uint32_t nq;
    uint64_t K;
    uint32_t nslices;  // value slices of a sparse query's bitmap
    uint64_t sp_cap;   // routed-record budget (0: the atomic path for every sparse query)
    // graph mode: the keys are copied to ctx->gkeys before every launch and
    // the key count is read on the device, so one graph serves every batch
    // of the same query count whose key count rounds up to the same Kcap
    bool graph;
    uint64_t Kcap;
    const uint32_t *keys;
    uint32_t G;
    int mode;
    bool use_col;
    uint32_t grid;  // work-groups of the persistent launch
    bool packed;    // per-query id offsets follow qoff in the staged block
    size_t stage_bytes;
    uint64_t mask_stride, any_stride;  // elements / u32 rank words per group
    uint32_t ngroups;                  // dense groups the batch can have
    uint32_t *bm;
    uint64_t stride, words;
};

// Every device launch of one batch; nothing here allocates or synchronizes,
// so it can be captured into a graph.
ab_status enqueue(ab_ctx *c, const BatchArgs &a, const Plan &p)
{
    ab_index *ix = c->ix;
    const DevIndex &d = ix->dev;
    hipStream_t s = c->stream;
    const uint32_t nq = p.nq;
    const uint64_t *qoff = (const uint64_t *)c->qoff.p;
    uint8_t *sel = (uint8_t *)c->sel.p;
    // the union's slice items count their queries' compaction chunks, so the
    // counting pass reads only the other queries' bitmaps
    const bool slice_counts = (a.d_ids || a.d_counts) && p.sp_cap && p.words;
    // qoff and, for packed ids, the id offsets arrive in one copy
    HIP_TRY(hipMemcpyAsync(c->qoff.p, c->hstage.p, p.stage_bytes, hipMemcpyHostToDevice, s));
    const uint64_t *Kd = p.graph ? qoff + nq : nullptr;  // qoff[nq] = the key count
    SparseArgs sa{};
    sa.keys = p.keys;
    sa.kq = (const uint32_t *)c->kq.p;
    sa.qoff = qoff;
    sa.koff = (const uint64_t *)c->koff.p;
    sa.K = p.K;
    sa.Kd = Kd;
    sa.nq = nq;
    sa.chunk = c->sparse_chunk;
    sa.nslices = p.nslices;
    sa.flag = (uint8_t *)c->sp_flag.p;
    sa.cpre = (uint32_t *)c->sp_cpre.p;
    sa.scpre = (uint32_t *)c->sp_scpre.p;
    sa.rbase = (uint64_t *)c->sp_rbase.p;
    sa.slist = (uint32_t *)c->sp_slist.p;
    sa.tot = (uint32_t *)c->sp_tot.p;
    sa.refs = (uint32_t *)c->sp_refs.p;
    sa.choff = (uint16_t *)c->sp_choff.p;
    sa.cap = p.sp_cap;
    sa.out = p.bm;
    sa.out_stride = p.stride;
    {
        TimedSpan ts(c, AB_K_DECIDE);
        // the per-query record counts M (features) split the sparse queries
        const uint64_t *M = nullptr;
        if (a.path == AB_PATH_DENSE || a.path == AB_PATH_SPARSE) {
            hipLaunchKernelGGL(k_fill_u8, dim3((nq + 255) / 256), dim3(256), 0, s, sel, nq, (uint8_t)a.path);
            if (a.path == AB_PATH_SPARSE) {
                launch_features(d, p.keys, qoff, nq, ix->R, (float *)c->feat.p, (uint64_t *)c->M.p, s);
                M = (const uint64_t *)c->M.p;
            }
        } else {
            launch_features(d, p.keys, qoff, nq, ix->R, (float *)c->feat.p, (uint64_t *)c->M.p, s);
            M = (const uint64_t *)c->M.p;
            const int slot = nq > 1 ? 1 : 0;
            if ((a.path == AB_PATH_LEARNED || a.path == AB_PATH_AUTO) && c->mlp[slot].loaded) {
                AB_TRY(mlp_decide(c, slot, (const float *)c->feat.p, nq, sel));
            } else {
                const double tau = nq > 1 ? c->tau_batch : c->tau;
                launch_decide_threshold((const uint64_t *)c->M.p, nq, tau * (double)ix->R, sel, s);
            }
        }
        launch_assign_groups(sel, M, nq, p.G, (uint32_t *)c->qgroup.p, (uint32_t *)c->grp_q.p, (uint32_t *)c->grp_n.p,
                             (uint32_t *)c->plan.p, sa, s);
    }
    if (!p.sp_cap) {
        // the atomic path ORs into zeroed bitmaps; slice queries write theirs whole
        TimedSpan ts(c, AB_K_SPARSE_ZERO);
        launch_zero_sparse(p.bm, p.stride, p.words, sa.flag, nq, kSpAtomic, s);
    }
    const uint32_t *kq = (const uint32_t *)c->kq.p;
    if (p.Kcap) {
        // the query of every key (kq), which the group masks use too, and the
        // sparse queries' per-key record counts
        TimedSpan ts(c, AB_K_SPARSE_PLAN);
        launch_sparse_plan(d, p.keys, qoff, sel, nq, (uint32_t *)c->kq.p, (uint64_t *)c->cnt.p, s);
    }
    {
        TimedSpan ts(c, AB_K_DENSE_MASK);
        launch_mask_groups(p.mode, p.keys, kq, Kd, p.Kcap, (const uint32_t *)c->qgroup.p, c->gmask.p,
                           p.mask_stride, (uint32_t *)c->ganybm.p, p.any_stride, ix->nkeys, kMaskAny, s);
        launch_mask_rank((uint32_t *)c->ganybm.p, p.any_stride, ix->nkeys, p.ngroups, (const uint32_t *)c->plan.p, s);
        launch_mask_groups(p.mode, p.keys, kq, Kd, p.Kcap, (const uint32_t *)c->qgroup.p, c->gmask.p,
                           p.mask_stride, (uint32_t *)c->ganybm.p, p.any_stride, ix->nkeys, kMaskSet, s);
    }
    if (p.Kcap) {
        {
            TimedSpan ts(c, AB_K_SPARSE_PLAN);
            launch_scan_u64((const uint64_t *)c->cnt.p, (uint64_t *)c->koff.p, p.Kcap, (uint64_t *)c->parts.p, s,
                            Kd);
        }
        TimedSpan ts(c, AB_K_SPARSE_UNION);
        HIP_TRY(launch_sparse_route(d, sa, c->sp_grid, s));
    }
    {
        TimedSpan ts(c, AB_K_FUSED_UNION);
        FusedArgs f{};
        {
            f.sp_tot = sa.tot;
            f.sp_slist = sa.slist;
            f.sp_cpre = sa.cpre;
            f.sp_scpre = sa.scpre;
            f.sp_refs = sa.refs;
            f.sp_rbase = sa.rbase;
            f.sp_choff = sa.choff;
            f.sp_chunk = sa.chunk;
            f.nslices = sa.nslices;
        }
        f.plan = (const uint32_t *)c->plan.p;
        f.grp_q = (const uint32_t *)c->grp_q.p;
        f.grp_n = (const uint32_t *)c->grp_n.p;
        f.masks = c->gmask.p;
        f.mask_stride = p.mask_stride;
        f.anybms = (const uint32_t *)c->ganybm.p;
        f.any_stride = p.any_stride;
        f.out = p.bm;
        f.out_stride = p.stride;
        f.out_words = p.words;
        f.use_col = p.use_col ? 1 : 0;
        f.counter = (uint32_t *)c->counter.p;
        f.ngroups_cap = p.ngroups;
        if (slice_counts) {
            f.ccnt = (uint32_t *)c->ccnt.p;
            f.nchunk = (uint32_t)((p.words + kCompactWords - 1) / kCompactWords);
        }
        HIP_TRY(launch_union_fused(p.mode, c->dense_threads, d, f, p.G, p.grid, s));
    }
    {
        TimedSpan ts(c, AB_K_DENSE_MASK);
        launch_mask_groups(p.mode, p.keys, kq, Kd, p.Kcap, (const uint32_t *)c->qgroup.p, c->gmask.p,
                           p.mask_stride, (uint32_t *)c->ganybm.p, p.any_stride, ix->nkeys, kMaskClear, s);
        launch_fill_zero(c->ganybm.p, (size_t)p.ngroups * p.any_stride * 4, s);
    }
    if (a.d_ids || a.d_counts) {
        {
            TimedSpan ts(c, AB_K_POPCOUNT);
            launch_popcount(p.bm, p.stride, p.words, nq, (uint32_t *)c->ccnt.p, s, slice_counts ? sa.flag : nullptr);
        }
        TimedSpan ts(c, AB_K_WRITE_IDS);
        const uint64_t *d_idoff = p.packed ? qoff + (nq + 1) : nullptr;
        launch_write_ids(p.bm, p.stride, p.words, nq, (uint32_t *)c->ccnt.p, a.d_ids, a.ids_stride, d_idoff,
                         a.id_off_aligned, a.d_counts, s);
    }
    HIP_TRY(hipGetLastError());
    return AB_OK;
}

}  // namespace

ab_status run_batch_fused(ab_ctx *c, const BatchArgs &a)
{
    ab_index *ix = c->ix;
    hipStream_t s = c->stream;
    Plan p{};
    p.nq = a.nq;
    p.K = a.h_qoff[a.nq];
    p.graph = c->batch_mode == 2 && !c->timing;
    p.Kcap = p.K;
    if (p.graph && p.K) {
        p.Kcap = 4096;
        while (p.Kcap < p.K) p.Kcap <<= 1;
    }
    // a group never needs to be wider than the batch: a narrower group takes
    // narrower masks (u32 up to 32 queries) and less LDS
    p.G = std::max<uint32_t>(1, std::min<uint32_t>(max_group(c), a.nq));
    p.mode = p.G <= 32 ? kModeMask32 : (p.G <= 64 ? kModeMask64 : (p.G <= 128 ? kModeMask128 : kModeMask256));
    p.use_col = col_fits(c, p.mode, p.G);
    HIP_TRY(fused_grid(p.mode, c->dense_threads, ix->dev, p.G, p.use_col, &p.grid));
    c->fused_grid = p.grid;
    p.stride = a.out_stride;
    p.words = a.out_words;
    const uint32_t nq = a.nq;
    const uint32_t ngmax = (nq + p.G - 1) / p.G;
    p.ngroups = ngmax;
    const uint64_t nkeys = ix->nkeys;
    p.mask_stride = (nkeys + 63) / 64 * 64;
    p.any_stride = mask_rank_words(nkeys);
    const size_t esize = p.mode == kModeMask32 ? 4 : (p.mode == kModeMask64 ? 8 : (p.mode == kModeMask128 ? 16 : 32));

    // write-once sparse: slices of 2^18 values; refs hold type << 29 | index
    p.nslices = (uint32_t)std::max<uint64_t>(1, ((uint64_t)ix->nvalues + (1ull << kSpSliceShift) - 1) >> kSpSliceShift);
    p.sp_cap = c->sp_cap;
    for (int t = 0; t < AB_NTYPES; ++t)
        if (ix->nrec[t] >= (1ull << kSpTypeShift)) p.sp_cap = 0;
    if (p.nslices > kSpMaxSlices) p.sp_cap = 0;

    // ---- every buffer exists before anything is enqueued (capture-safe) ----
    const uint64_t nchunk = (p.words + kCompactWords - 1) / kCompactWords;
    std::vector<DevBuf *> bufs = {&c->qoff, &c->sel, &c->feat, &c->M, &c->qgroup, &c->grp_q, &c->grp_n, &c->plan,
                                  &c->counter, &c->kq, &c->cnt, &c->koff, &c->parts, &c->ccnt, &c->bm, &c->gkeys,
                                  &c->sp_flag, &c->sp_cpre, &c->sp_scpre, &c->sp_rbase, &c->sp_slist, &c->sp_tot,
                                  &c->sp_refs, &c->sp_choff};
    std::vector<void *> before;
    for (DevBuf *b : bufs) before.push_back(b->p);
    p.packed = a.h_id_off && a.d_ids;
    p.stage_bytes = ((size_t)nq + 1) * 8 + (p.packed ? (size_t)nq * 8 : 0);
    AB_TRY(c->qoff.ensure(p.stage_bytes));
    AB_TRY(c->sel.ensure(nq));
    AB_TRY(c->feat.ensure((size_t)nq * AB_NFEATURES * 4));
    AB_TRY(c->M.ensure((size_t)nq * 8));
    AB_TRY(c->qgroup.ensure((size_t)nq * 4));
    AB_TRY(c->grp_q.ensure((size_t)ngmax * kGroupSlots * 4));
    AB_TRY(c->grp_n.ensure((size_t)ngmax * 4 + 4));
    AB_TRY(c->plan.ensure(16));
    AB_TRY(c->counter.ensure(((size_t)ngmax + 2) * 4));  // slice items, then one per dense group
    AB_TRY(c->kq.ensure(p.Kcap * 4 + 4));
    AB_TRY(c->cnt.ensure(p.Kcap * 8 + 8));
    AB_TRY(c->koff.ensure((p.Kcap + 1) * 8));
    AB_TRY(c->parts.ensure(scan_parts_needed(p.Kcap) * 8 + 16));
    if (p.graph) AB_TRY(c->gkeys.ensure(p.Kcap * 4 + 16));
    AB_TRY(c->ccnt.ensure((size_t)nq * nchunk * 4 + 4));
    {
        AB_TRY(c->sp_flag.ensure(nq + 16));
        AB_TRY(c->sp_cpre.ensure(((size_t)nq + 1) * 4));
        AB_TRY(c->sp_scpre.ensure((size_t)nq * 4 + 4));
        AB_TRY(c->sp_rbase.ensure((size_t)nq * 8 + 8));
        AB_TRY(c->sp_slist.ensure((size_t)nq * 4 + 4));
        AB_TRY(c->sp_tot.ensure(16));
        if (p.sp_cap) {
            // a slice query of M records has ceil(M / chunk) <= M / chunk + 1 chunks
            const uint64_t max_chunks = p.sp_cap / c->sparse_chunk + nq + 1;
            AB_TRY(c->sp_refs.ensure(p.sp_cap * 4));
            AB_TRY(c->sp_choff.ensure(max_chunks * (p.nslices + 1) * 2));
        }
        if (!c->sp_grid) {
            int dev = 0, cus = 0, per = 0;
            HIP_TRY(hipGetDevice(&dev));
            HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            per = 4;
            c->sp_grid = (uint32_t)(cus * per);
        }
    }
    p.bm = a.d_bitmaps;
    if (!p.bm) {
        AB_TRY(c->bm.ensure((size_t)nq * p.stride * 4));
        p.bm = (uint32_t *)c->bm.p;
    }
    if (c->gmask_groups < ngmax || c->gmask_mode != p.mode) {
        // mask buffers start zeroed and are returned to zero by every batch
        AB_TRY(c->gmask.ensure((size_t)ngmax * p.mask_stride * esize));
        AB_TRY(c->ganybm.ensure((size_t)ngmax * p.any_stride * 4));
        // (stream-ordered, never inside a capture: this runs before any)
        HIP_TRY(hipMemsetAsync(c->gmask.p, 0, c->gmask.cap, s));
        HIP_TRY(hipMemsetAsync(c->ganybm.p, 0, c->ganybm.cap, s));
        c->gmask_groups = (uint32_t)std::min<size_t>(c->gmask.cap / (p.mask_stride * esize),
                                                     c->ganybm.cap / (p.any_stride * 4));
        c->gmask_mode = p.mode;
        release_graph(c);
    }
    std::vector<void *> after;
    for (DevBuf *b : bufs) after.push_back(b->p);
    if (before != after) release_graph(c);
    p.keys = p.graph ? (const uint32_t *)c->gkeys.p : a.d_keys;

    // ---- stage the offsets (the graph's copy node reads this pinned area) ----
    AB_TRY(stage_wait(c));
    AB_TRY(c->hstage.ensure(p.stage_bytes + 64));
    memcpy(c->hstage.p, a.h_qoff, ((size_t)nq + 1) * 8);
    if (p.packed) memcpy((uint8_t *)c->hstage.p + ((size_t)nq + 1) * 8, a.h_id_off, (size_t)nq * 8);

    if (!p.graph) {
        AB_TRY(enqueue(c, a, p));
    } else {
        // this batch's keys into the buffer the graph reads
        if (p.K) HIP_TRY(hipMemcpyAsync(c->gkeys.p, a.d_keys, p.K * 4, hipMemcpyDeviceToDevice, s));
        std::vector<uint64_t> key = {nq, p.Kcap, (uint64_t)a.path, p.grid, p.sp_cap, c->sparse_chunk,
                                     (uint64_t)(uintptr_t)p.bm, p.stride, p.words, (uint64_t)(uintptr_t)a.d_ids,
                                     a.ids_stride, (uint64_t)p.packed, (uint64_t)a.id_off_aligned,
                                     (uint64_t)(uintptr_t)a.d_counts, c->cfg_gen,
                                     (uint64_t)(uintptr_t)c->hstage.p, (uint64_t)(uintptr_t)c->gmask.p,
                                     (uint64_t)(uintptr_t)c->ganybm.p};
        for (void *q : after) key.push_back((uint64_t)(uintptr_t)q);
        const bool hit = c->gc.valid && c->gc.key == key;
        if (!hit && c->gc.pending != key) {
            // a shape seen once runs as launches; it is captured when it comes
            // again in the next call
            c->gc.pending = key;
            AB_TRY(enqueue(c, a, p));
        } else if (!hit) {
            release_graph(c);
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            ab_status st = enqueue(c, a, p);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (st != AB_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            HIP_TRY(e);
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                HIP_TRY(e);
            }
            c->gc.graph = g;
            c->gc.exec = ex;
            c->gc.key = key;
            c->gc.valid = true;
            HIP_TRY(hipGraphLaunch(c->gc.exec, s));
        } else {
            HIP_TRY(hipGraphLaunch(c->gc.exec, s));
        }
    }
    HIP_TRY(hipEventRecord(c->stage_ev, s));
    c->stage_pending = true;
    if (a.h_paths) {
        AB_TRY(c->hsel.ensure(nq));
        HIP_TRY(hipMemcpyAsync(c->hsel.p, c->sel.p, nq, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        memcpy(a.h_paths, c->hsel.p, nq);
    }
    return AB_OK;
}

}  // namespace ab
