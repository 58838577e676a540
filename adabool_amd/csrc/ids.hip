// ids.hip -- id-only batches: every query's ascending ids formed without a
// result bitmap.  Memory and traffic follow the values that the query's lists
// hold, not nvalues, so this serves indices whose per-query bitmaps would not
// fit (nvalues up to 2^32 - 1) and queries whose results are a small part of
// nvalues.
//
//   1. plan, one work-group per query: each key's record count from seg8
//      (the Prepare(keys) extents, src/AdaBool/inverted_list.h:491-505), the
//      query's record and value totals; the totals go back to the host, which
//      splits the batch into parts that fit the expansion budget;
//   2. per part: an exclusive scan numbers the records; work-groups take
//      chunks of kIdsChunk records, count their values (record_count), and a
//      scan over chunks places each chunk's values;
//   3. the same chunks again: a work-group scan of the record counts places
//      every record, whose values go out as (query << 32 | value);
//   4. radix sort of the packed values (rocPRIM);
//   5. first copies flagged and ranked by a scan; the first index of every
//      query's run; each first copy to the caller's offset for its query, and
//      the number of first copies as its count.
// A second method forms the ids with no sort, by value slice (for value
// spaces of at most kSpMaxSlices slices of 2^18 values and record indices
// that fit the 29-bit references):
//   2'. per part: the records are numbered as above and routed by value
//       slice into 4-byte references (the write-once sparse path's route,
//       every query a slice query);
//   3'. one persistent launch takes the (query, slice) items in order: the
//       slice's references OR-ed into a 32 KB LDS bitmap, its bits counted,
//       the query's ids before the slice found by a decoupled look-back over
//       the query's earlier slices, and the bitmap expanded straight to ids.
// Every item scans its whole slice, so the slices cost nq x S x 8192 words
// whatever the result, and the sort costs follow the values: the automatic
// choice takes the slices when the batch has at most kIdsSliceItems items or
// its values fill at least 1/kIdsSliceDensity of the slices' value space
// (measured on the C2 filter: 64-key queries are faster by slice up to batch
// 256 and by sort from batch 1024; C3-mix queries faster by slice at any batch).
// The result equals the bitmap paths' ids for every query: the set union of
// its in-range keys' lists, ascending (test/adsfilter_q_all.cu:108-119,
// 196-208).
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "device.h"
#include "engine.h"
#include "layout.h"

using namespace ab;

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

namespace {

constexpr uint32_t kIdsNT = 256;      // threads per work-group
constexpr uint32_t kIdsChunk = 4096;  // records per chunk
constexpr uint32_t kIdsWin = 1024;    // key offsets per LDS window
constexpr uint32_t kIdsGrid = 8192;   // work-groups of the chunk kernels (grid-stride)
constexpr uint32_t kIdsStage = 16384; // LDS bytes staging one round's values
// automatic method: by value slice up to this many (query, slice) items, or
// when the values fill at least 1/kIdsSliceDensity of the slices
constexpr uint64_t kIdsSliceItems = 8192;
constexpr uint64_t kIdsSliceDensity = 20;
// records per part of the slice method (2 GiB of references)
constexpr uint64_t kIdsSliceRecords = 1ull << 29;

// Exclusive work-group scan of one u32 per thread; *total = the sum.
__device__ __forceinline__ uint32_t block_excl_u32(uint32_t x, uint32_t *sw, uint32_t *total)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(x);
    if (lane == 63) sw[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kIdsNT / 64; ++w) {
        before += w < wave ? sw[w] : 0u;
        all += sw[w];
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

// Per query q: rcnt[j] = records of key j (0 when out of range), kq[j] = q,
// tot[2q] = sum of rcnt over the query, tot[2q+1] = its values (list lengths
// with repeats).
__global__ __launch_bounds__(kIdsNT) void k_ids_plan(DevIndex ix, const uint32_t *__restrict__ keys,
                                                     const uint64_t *__restrict__ qoff, uint32_t *__restrict__ kq,
                                                     uint64_t *__restrict__ rcnt, uint64_t *__restrict__ tot)
{
    __shared__ unsigned long long sr[kIdsNT / 64], sv[kIdsNT / 64];
    const uint32_t q = blockIdx.x, lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t a = qoff[q], b = qoff[q + 1];
    unsigned long long r = 0, v = 0;
    for (uint64_t j = a + threadIdx.x; j < b; j += kIdsNT) {
        const uint32_t k = keys[j];
        uint64_t c = 0;
        if (k < ix.nkeys) {
            const uint32_t *lo = ix.seg8 + (uint64_t)k * kSegStride, *hi = lo + kSegStride;
            c = hi[6] - lo[6];
            v += hi[7] - lo[7];
        }
        rcnt[j] = c;
        kq[j] = q;
        r += c;
    }
    r = wave_sum64(r);
    v = wave_sum64(v);
    if (lane == 0) {
        sr[wave] = r;
        sv[wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long R = 0, V = 0;
        for (uint32_t w = 0; w < kIdsNT / 64; ++w) {
            R += sr[w];
            V += sv[w];
        }
        tot[2ull * q] = R;
        tot[2ull * q + 1] = V;
    }
}

// One record, located by key j's seg8 rows and its rank r among the key's
// records (types in order ARRAY1..BITSET).
struct Rec {
    int t;
    uint32_t h;
    uint32_t w[8];
};

template <int T>
__device__ __forceinline__ void load_rec(const DevIndex &ix, uint32_t idx, Rec &rc)
{
    rc.t = T;
    rc.h = ix.hdr[T][idx];
#pragma unroll
    for (int i = 0; i < 8; ++i) rc.w[i] = 0;
    load_payload<T>(ix.pay[T], idx, rc.w);
}

__device__ __forceinline__ void locate(const DevIndex &ix, uint32_t k, uint32_t r, Rec &rc)
{
    const uint4 *rows = reinterpret_cast<const uint4 *>(ix.seg8 + (uint64_t)k * kSegStride);
    const uint4 a0 = rows[0], a1 = rows[1], b0 = rows[2], b1 = rows[3];
    const uint32_t l0 = b0.x - a0.x, l1 = b0.y - a0.y, l2 = b0.z - a0.z, l3 = b0.w - a0.w, l4 = b1.x - a1.x;
    if (r < l0) return load_rec<AB_ARRAY1>(ix, a0.x + r, rc);
    r -= l0;
    if (r < l1) return load_rec<AB_ARRAY4>(ix, a0.y + r, rc);
    r -= l1;
    if (r < l2) return load_rec<AB_ARRAY8>(ix, a0.z + r, rc);
    r -= l2;
    if (r < l3) return load_rec<AB_ARRAY16>(ix, a0.w + r, rc);
    r -= l3;
    if (r < l4) return load_rec<AB_ARRAY32>(ix, a1.x + r, rc);
    r -= l4;
    load_rec<AB_BITSET>(ix, a1.y + r, rc);
}

__device__ __forceinline__ uint32_t rec_count(const Rec &rc)
{
    switch (rc.t) {
    case AB_ARRAY1: return record_count<AB_ARRAY1>(rc.w);
    case AB_ARRAY4: return record_count<AB_ARRAY4>(rc.w);
    case AB_ARRAY8: return record_count<AB_ARRAY8>(rc.w);
    case AB_ARRAY16: return record_count<AB_ARRAY16>(rc.w);
    case AB_ARRAY32: return record_count<AB_ARRAY32>(rc.w);
    default: return record_count<AB_BITSET>(rc.w);
    }
}

template <class F>
__device__ __forceinline__ void rec_values(const Rec &rc, F &&f)
{
    switch (rc.t) {
    case AB_ARRAY1: record_values<AB_ARRAY1>(rc.h, rc.w, f); break;
    case AB_ARRAY4: record_values<AB_ARRAY4>(rc.h, rc.w, f); break;
    case AB_ARRAY8: record_values<AB_ARRAY8>(rc.h, rc.w, f); break;
    case AB_ARRAY16: record_values<AB_ARRAY16>(rc.h, rc.w, f); break;
    case AB_ARRAY32: record_values<AB_ARRAY32>(rc.h, rc.w, f); break;
    default: record_values<AB_BITSET>(rc.h, rc.w, f); break;
    }
}

// Records [c0, c1) of a part, in rounds of kIdsNT (one record per thread).
// koff: the part's record prefix over its K keys.  EMIT = false: returns the
// chunk's value count; EMIT = true: writes the chunk's values from out[base]
// on, in record order, as (query - q0) << vbits | value (never at or past
// limit).  KT: u32 when query and value bits fit, else u64.
template <bool EMIT, class KT>
__device__ uint64_t ids_chunk(const DevIndex &ix, const uint32_t *__restrict__ keys,
                              const uint32_t *__restrict__ kq, const uint64_t *__restrict__ koff, uint64_t K,
                              uint64_t c0, uint64_t c1, uint32_t q0, uint32_t vbits, KT *__restrict__ out,
                              uint64_t base, uint64_t limit, KT *stage, unsigned long long *skoff,
                              unsigned long long *s_lo, unsigned long long *s_hi, uint32_t *sw)
{
    uint64_t jw = wg_find<kIdsNT>(koff, K, c0, s_lo, s_hi);
    uint64_t run = 0;
    for (;;) {
        const uint64_t jn = (jw + kIdsWin < K) ? jw + kIdsWin : K;
        const uint32_t nwin = (uint32_t)(jn - jw);
        for (uint32_t i = threadIdx.x; i <= nwin; i += kIdsNT) skoff[i] = koff[jw + i];
        __syncthreads();
        const uint64_t wlo = skoff[0] > c0 ? skoff[0] : c0;
        const uint64_t whi = skoff[nwin] < c1 ? skoff[nwin] : c1;
        for (uint64_t it0 = wlo; it0 < whi; it0 += kIdsNT) {
            const uint64_t it = it0 + threadIdx.x;
            Rec rc;
            uint32_t cnt = 0, q = 0;
            if (it < whi) {
                uint32_t lo = 0, hi = nwin;  // skoff[lo] <= it < skoff[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (skoff[mid] <= it) lo = mid; else hi = mid;
                }
                const uint64_t j = jw + lo;
                q = kq[j] - q0;
                locate(ix, keys[j], (uint32_t)(it - skoff[lo]), rc);
                cnt = rec_count(rc);
            }
            uint32_t tot;
            const uint32_t ex = block_excl_u32(cnt, sw, &tot);
            if constexpr (EMIT) {
                const KT qk = (KT)q << vbits;
                if (tot <= kIdsStage / sizeof(KT)) {
                    // through LDS, so that the round's values leave in
                    // consecutive, coalesced stores
                    if (it < whi) {
                        uint32_t p = ex;
                        rec_values(rc, [&](uint32_t v) { stage[p++] = qk | (KT)v; });
                    }
                    __syncthreads();
                    for (uint32_t i = threadIdx.x; i < tot; i += kIdsNT) {
                        const uint64_t pos = base + run + i;
                        if (pos < limit) out[pos] = stage[i];
                    }
                    __syncthreads();
                } else if (it < whi) {
                    uint64_t pos = base + run + ex;
                    rec_values(rc, [&](uint32_t v) {
                        if (pos < limit) out[pos] = qk | (KT)v;
                        ++pos;
                    });
                }
            }
            run += tot;
        }
        const bool done = (skoff[nwin] >= c1) || jn >= K;
        __syncthreads();
        if (done) break;
        jw = jn;
    }
    return run;
}

__global__ __launch_bounds__(kIdsNT) void k_ids_count(DevIndex ix, const uint32_t *__restrict__ keys,
                                                      const uint32_t *__restrict__ kq,
                                                      const uint64_t *__restrict__ koff, uint64_t K, uint64_t W,
                                                      uint64_t *__restrict__ ccnt)
{
    __shared__ unsigned long long skoff[kIdsWin + 1];
    __shared__ unsigned long long s_lo, s_hi;
    __shared__ uint32_t sw[kIdsNT / 64];
    const uint64_t nch = (W + kIdsChunk - 1) / kIdsChunk;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint64_t c0 = c * kIdsChunk, c1 = (c0 + kIdsChunk < W) ? c0 + kIdsChunk : W;
        const uint64_t n = ids_chunk<false, uint32_t>(ix, keys, kq, koff, K, c0, c1, 0, 0, nullptr, 0, 0, nullptr, skoff,
                                                     &s_lo, &s_hi, sw);
        if (threadIdx.x == 0) ccnt[c] = n;
    }
}

template <class KT>
__global__ __launch_bounds__(kIdsNT) void k_ids_expand(DevIndex ix, const uint32_t *__restrict__ keys,
                                                       const uint32_t *__restrict__ kq,
                                                       const uint64_t *__restrict__ koff, uint64_t K, uint64_t W,
                                                       const uint64_t *__restrict__ cbase, uint32_t q0,
                                                       uint32_t vbits, KT *__restrict__ out, uint64_t limit)
{
    __shared__ unsigned long long skoff[kIdsWin + 1];
    __shared__ unsigned long long s_lo, s_hi;
    __shared__ uint32_t sw[kIdsNT / 64];
    __shared__ __attribute__((aligned(16))) KT stage[kIdsStage / sizeof(KT)];
    const uint64_t nch = (W + kIdsChunk - 1) / kIdsChunk;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint64_t c0 = c * kIdsChunk, c1 = (c0 + kIdsChunk < W) ? c0 + kIdsChunk : W;
        (void)ids_chunk<true, KT>(ix, keys, kq, koff, K, c0, c1, q0, vbits, out, cbase[c], limit, stage, skoff,
                                   &s_lo, &s_hi, sw);
    }
}

__global__ void k_ids_zero(uint32_t *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// f[i] = 1 where the sorted packed value x[i] differs from x[i - 1]: the
// first copy of a value within its query.
template <class KT>
__global__ void k_ids_flags(const KT *__restrict__ x, uint64_t n, uint64_t *__restrict__ f)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        f[i] = (i == 0 || x[i] != x[i - 1]) ? 1u : 0u;
}

// start[q] = index of query q's first element in the sorted array x.
template <class KT>
__global__ void k_ids_starts(const KT *__restrict__ x, uint64_t n, uint32_t vbits, uint64_t *__restrict__ start)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(x[i] >> vbits);
        if (i == 0 || (uint32_t)(x[i - 1] >> vbits) != q) start[q] = i;
    }
}

// ex: exclusive scan of the first-copy flags.  A first copy goes to
// ids[id_off[q0 + q] + ex[i] - ex[start[q]]]; the last element of a query's
// run writes its count.
template <class KT>
__global__ void k_ids_write(const KT *__restrict__ x, uint64_t n, uint32_t vbits, const uint64_t *__restrict__ ex,
                            const uint64_t *__restrict__ start, const uint64_t *__restrict__ id_off, uint32_t q0,
                            uint32_t *__restrict__ ids, uint32_t *__restrict__ counts)
{
    const KT vmask = (KT)(((unsigned long long)1 << vbits) - 1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const KT v = x[i];
        const uint32_t q = (uint32_t)(v >> vbits);
        const uint64_t base = ex[start[q]];
        const bool first = i == 0 || v != x[i - 1];
        if (first) ids[id_off[q0 + q] + (ex[i] - base)] = (uint32_t)(v & vmask);
        if (i + 1 == n || (uint32_t)(x[i + 1] >> vbits) != q) counts[q0 + q] = (uint32_t)(ex[i] + first - base);
    }
}

int bits_for(uint64_t x)
{
    int b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b;
}

// Expansion, sort and write of one part whose chunk value counts are scanned
// into id_cbase.
template <class KT>
ab_status sort_and_write(ab_ctx *c, const DevIndex &d, const uint32_t *keys_p, const uint32_t *kq_p, uint64_t Kp,
                         uint64_t W, uint32_t grid, uint32_t q0, int vbits, int end_bit, uint64_t T, uint32_t *d_ids,
                         uint32_t *counts)
{
    hipStream_t s = c->stream;
    KT *vals = (KT *)c->id_vals.p, *srt = (KT *)c->id_sorted.p;
    hipLaunchKernelGGL(k_ids_expand<KT>, dim3(grid), dim3(kIdsNT), 0, s, d, keys_p, kq_p,
                       (const uint64_t *)c->koff.p, Kp, W, (const uint64_t *)c->id_cbase.p, q0, (uint32_t)vbits, vals,
                       T);
    HIP_TRY(hipGetLastError());
    size_t tb = 0;
    HIP_TRY(rocprim::radix_sort_keys(nullptr, tb, vals, srt, (size_t)T, 0, end_bit, s));
    AB_TRY(c->id_tmp.ensure(tb + 64));
    HIP_TRY(rocprim::radix_sort_keys(c->id_tmp.p, tb, vals, srt, (size_t)T, 0, end_bit, s));
    // first copies flagged, then ranked by a scan (the expansion buffer holds
    // the flags, then their exclusive scan)
    const uint32_t wgrid = (uint32_t)std::min<uint64_t>((T + 255) / 256, 16384);
    uint64_t *ex = (uint64_t *)c->id_vals.p;
    hipLaunchKernelGGL(k_ids_flags<KT>, dim3(wgrid), dim3(256), 0, s, srt, T, ex);
    launch_scan_u64(ex, ex, T, (uint64_t *)c->parts.p, s);
    hipLaunchKernelGGL(k_ids_starts<KT>, dim3(wgrid), dim3(256), 0, s, srt, T, (uint32_t)vbits,
                       (uint64_t *)c->id_start.p);
    hipLaunchKernelGGL(k_ids_write<KT>, dim3(wgrid), dim3(256), 0, s, srt, T, (uint32_t)vbits, (const uint64_t *)ex,
                       (const uint64_t *)c->id_start.p, (const uint64_t *)c->idoff.p, q0, d_ids, counts);
    HIP_TRY(hipGetLastError());
    return AB_OK;
}

}  // namespace

uint64_t ab::id_upper_bound(const ab_index *ix, const uint32_t *keys, uint64_t n)
{
    uint64_t ub = 0;
    const uint32_t *seg = ix->seg8.data();
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t k = keys[i];
        if (k >= ix->nkeys) continue;
        ub += seg[((uint64_t)k + 1) * kSegStride + 7] - seg[(uint64_t)k * kSegStride + 7];
        if (ub >= ix->nvalues) return ix->nvalues;
    }
    return ub;
}

extern "C" ab_status ab_query_id_bound(const ab_index *ix, const uint32_t *h_keys, uint32_t n, uint64_t *bound)
{
    if (!ix || !bound || (n && !h_keys)) {
        set_error("ab_query_id_bound: bad argument");
        return AB_INVALID_ARG;
    }
    *bound = id_upper_bound(ix, h_keys, n);
    return AB_OK;
}

extern "C" ab_status ab_ctx_set_id_budget(ab_ctx *c, uint64_t values)
{
    if (!c || values < kIdsChunk) {
        set_error("ab_ctx_set_id_budget: budget below one chunk");
        return AB_INVALID_ARG;
    }
    c->id_budget = values;
    return AB_OK;
}

namespace {

// Id-only parts by value slice (see the file comment): per part, route every
// query's records into slice-sorted references, then one persistent launch of
// the (query, slice) items.  tot: per query {records, values} from the plan.
ab_status ids_by_slice(ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq, uint32_t *d_ids,
                       const uint64_t *h_id_off, uint32_t *counts, const std::vector<uint64_t> &tot,
                       const std::vector<uint32_t> &part_q0, const std::vector<uint32_t> &part_q1, uint32_t S)
{
    ab_index *ix = c->ix;
    const DevIndex &d = ix->dev;
    hipStream_t s = c->stream;
    const uint32_t chunk = c->sparse_chunk;
    // ids + id_off[q] 16-byte aligned for every query: 128-bit stores
    bool aligned = ((uintptr_t)d_ids % 16) == 0;
    for (uint32_t q = 0; q < nq && aligned; ++q) aligned = h_id_off[q] % 4 == 0;
    if (!c->sp_grid) {
        int dev = 0, cus = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        c->sp_grid = (uint32_t)(cus * 4);
    }
    if (!c->ids_grid) c->ids_grid = ids_slices_grid();
    // per-part route arrays, all parts staged at once (distinct pinned
    // regions, so no copy of an earlier part is overwritten): qoff relative
    // to the part's first key, refs bases, route-chunk prefix, totals, flags
    auto a8 = [](size_t x) { return (x + 7) & ~(size_t)7; };
    std::vector<size_t> at(part_q0.size() + 1, 0);
    for (size_t pi = 0; pi < part_q0.size(); ++pi) {
        const size_t n = part_q1[pi] - part_q0[pi];
        at[pi + 1] = at[pi] + a8((n + 1) * 8) + a8(n * 8) + a8((n + 1) * 4) + 16 + a8(n);
    }
    AB_TRY(c->hidpart.ensure(at.back() + 64));
    size_t dev_bytes = 0;
    for (size_t pi = 0; pi < part_q0.size(); ++pi) dev_bytes = std::max(dev_bytes, at[pi + 1] - at[pi]);
    AB_TRY(c->idpart.ensure(dev_bytes + 64));
    for (size_t pi = 0; pi < part_q0.size(); ++pi) {
        const uint32_t q0 = part_q0[pi], q1 = part_q1[pi], n = q1 - q0;
        const uint64_t k0 = h_qoff[q0], Kp = h_qoff[q1] - k0;
        uint8_t *h = (uint8_t *)c->hidpart.p + at[pi];
        uint64_t *qoffp = (uint64_t *)h;
        uint64_t *rbase = (uint64_t *)(h + a8((n + 1) * 8));
        uint32_t *cpre = (uint32_t *)(h + a8((n + 1) * 8) + a8(n * 8));
        uint32_t *ptot = cpre + ((a8((n + 1) * 4)) / 4);
        uint8_t *flag = (uint8_t *)(ptot + 4);
        uint64_t W = 0, nch = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t M = tot[2ull * (q0 + i)];
            qoffp[i] = h_qoff[q0 + i] - k0;
            rbase[i] = W;
            cpre[i] = (uint32_t)nch;
            flag[i] = kSpSlice;
            W += M;
            nch += (M + chunk - 1) / chunk;
        }
        qoffp[n] = Kp;
        cpre[n] = (uint32_t)nch;
        ptot[0] = (uint32_t)nch;
        ptot[1] = n;
        ptot[2] = (uint32_t)nch;
        ptot[3] = 0;
        if (nch >= (1ull << 32) || W >= (1ull << 32)) {
            set_error("ab_query_batch_ids: part too large for the slice route");
            return AB_UNSUPPORTED;
        }
        uint8_t *dp = (uint8_t *)c->idpart.p;
        HIP_TRY(hipMemcpyAsync(dp, h, at[pi + 1] - at[pi], hipMemcpyHostToDevice, s));
        // record numbering over the part's keys
        AB_TRY(c->koff.ensure((Kp + 1) * 8));
        AB_TRY(c->parts.ensure((scan_parts_needed(Kp) + 2) * 8));
        launch_scan_u64((const uint64_t *)c->cnt.p + k0, (uint64_t *)c->koff.p, Kp, (uint64_t *)c->parts.p, s);
        AB_TRY(c->sp_refs.ensure(W * 4 + 16));
        AB_TRY(c->sp_choff.ensure(nch * (S + 1) * 2 + 16));
        AB_TRY(c->id_ccnt.ensure((size_t)n * S * 8 + 64));
        SparseArgs sa{};
        sa.keys = d_keys + k0;
        sa.kq = (const uint32_t *)c->kq.p + k0;
        sa.qoff = (const uint64_t *)dp;
        sa.koff = (const uint64_t *)c->koff.p;
        sa.K = Kp;
        sa.nq = n;
        sa.chunk = chunk;
        sa.nslices = S;
        sa.flag = dp + ((uint8_t *)flag - h);
        sa.cpre = (uint32_t *)(dp + ((uint8_t *)cpre - h));
        sa.scpre = sa.cpre;
        sa.rbase = (uint64_t *)(dp + ((uint8_t *)rbase - h));
        sa.tot = (uint32_t *)(dp + ((uint8_t *)ptot - h));
        sa.refs = (uint32_t *)c->sp_refs.p;
        sa.choff = (uint16_t *)c->sp_choff.p;
        sa.cap = W ? W : 1;
        HIP_TRY(launch_sparse_route(d, sa, c->sp_grid, s));
        // status words, work counter and error word start at zero
        unsigned long long *st = (unsigned long long *)c->id_ccnt.p;
        uint32_t *counter = (uint32_t *)(st + (size_t)n * S);
        launch_fill_zero(st, (size_t)n * S * 8 + 16, s);
        IdsSliceArgs ia{};
        ia.cpre = sa.cpre;
        ia.rbase = sa.rbase;
        ia.refs = sa.refs;
        ia.choff = sa.choff;
        ia.chunk = chunk;
        ia.nslices = S;
        ia.nq = n;
        ia.q0 = q0;
        ia.id_off = (const uint64_t *)c->idoff.p;
        ia.ids = d_ids;
        ia.counts = counts;
        ia.st = st;
        ia.counter = counter;
        ia.err = counter + 1;
        HIP_TRY(launch_ids_slices(d, ia, aligned, c->ids_grid, s));
        // the next part reuses koff, refs, choff and the status words
    }
    return AB_OK;
}

}  // namespace

extern "C" ab_status ab_ctx_set_id_method(ab_ctx *c, int method)
{
    if (!c || method < 0 || method > 2) {
        set_error("ab_ctx_set_id_method: 0 (automatic), 1 (by sort) or 2 (by value slice where it applies)");
        return AB_INVALID_ARG;
    }
    c->id_method = method;
    return AB_OK;
}

extern "C" ab_status ab_query_batch_ids(ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq,
                                        uint32_t *d_ids, const uint64_t *h_id_off, uint32_t *d_counts)
{
    if (!c || !h_qoff || !d_ids || !h_id_off || (nq && h_qoff[0] != 0) || c->ix->host_only) {
        set_error("ab_query_batch_ids: bad argument");
        return AB_INVALID_ARG;
    }
    if (nq == 0) return AB_OK;
    for (uint32_t q = 0; q < nq; ++q)
        if (h_qoff[q + 1] < h_qoff[q]) {
            set_error("ab_query_batch_ids: offsets must not decrease");
            return AB_INVALID_ARG;
        }
    const uint64_t K = h_qoff[nq];
    if (K && !d_keys) {
        set_error("ab_query_batch_ids: NULL keys");
        return AB_INVALID_ARG;
    }
    ab_index *ix = c->ix;
    const DevIndex &d = ix->dev;
    hipStream_t s = c->stream;
    TimedSpan span(c, AB_K_IDS);

    // ---- plan: per-key record counts, per-query totals back to the host ------
    AB_TRY(stage_wait(c));
    const size_t qoff_bytes = ((size_t)nq + 1) * 8, idoff_bytes = (size_t)nq * 8;
    AB_TRY(c->hstage.ensure(qoff_bytes + idoff_bytes + 64));
    memcpy(c->hstage.p, h_qoff, qoff_bytes);
    memcpy((uint8_t *)c->hstage.p + qoff_bytes, h_id_off, idoff_bytes);
    AB_TRY(c->qoff.ensure(qoff_bytes));
    AB_TRY(c->idoff.ensure(idoff_bytes));
    HIP_TRY(hipMemcpyAsync(c->qoff.p, c->hstage.p, qoff_bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->idoff.p, (uint8_t *)c->hstage.p + qoff_bytes, idoff_bytes, hipMemcpyHostToDevice, s));
    AB_TRY(c->kq.ensure(K * 4 + 4));
    AB_TRY(c->cnt.ensure(K * 8 + 8));
    AB_TRY(c->M.ensure((size_t)nq * 16));
    AB_TRY(c->hM.ensure((size_t)nq * 16));
    hipLaunchKernelGGL(k_ids_plan, dim3(nq), dim3(kIdsNT), 0, s, d, d_keys, (const uint64_t *)c->qoff.p,
                       (uint32_t *)c->kq.p, (uint64_t *)c->cnt.p, (uint64_t *)c->M.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(c->hM.p, c->M.p, (size_t)nq * 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<uint64_t> tot((const uint64_t *)c->hM.p, (const uint64_t *)c->hM.p + 2ull * nq);

    uint32_t *counts = d_counts;
    if (!counts) {
        AB_TRY(c->ccnt.ensure((size_t)nq * 4 + 4));
        counts = (uint32_t *)c->ccnt.p;
    }
    hipLaunchKernelGGL(k_ids_zero, dim3((nq + 255) / 256), dim3(256), 0, s, counts, nq);

    // ---- the method ---------------------------------------------------------
    const uint64_t S = std::max<uint64_t>(1, ((uint64_t)ix->nvalues + (1ull << kSpSliceShift) - 1) >> kSpSliceShift);
    bool by_slice = S <= kSpMaxSlices && c->id_method != 1;
    for (int t = 0; t < AB_NTYPES; ++t)
        if (ix->nrec[t] >= (1ull << kSpTypeShift)) by_slice = false;
    if (by_slice && c->id_method == 0) {
        uint64_t T = 0;
        for (uint32_t q = 0; q < nq; ++q) T += tot[2ull * q + 1];
        const uint64_t items = (uint64_t)nq * S;
        by_slice = items <= kIdsSliceItems || T * kIdsSliceDensity >= items << kSpSliceShift;
    }

    // ---- parts: by sort, of at most id_budget values (8 bytes each, twice);
    // by slice, of at most 4 x id_budget records (4-byte references, the same
    // scratch), up to kIdsSliceRecords; a larger query forms a part of its own
    const uint64_t budget = by_slice ? std::min<uint64_t>(kIdsSliceRecords, 4 * c->id_budget) : c->id_budget;
    const int what = by_slice ? 0 : 1;  // tot[2q]: records, tot[2q + 1]: values
    const int vbits = std::max(1, bits_for(ix->nvalues ? ix->nvalues - 1 : 0));
    std::vector<uint32_t> part_q0, part_q1;
    for (uint32_t q0 = 0; q0 < nq;) {
        uint64_t T = tot[2ull * q0 + what], V = tot[2ull * q0 + 1];
        uint32_t q1 = q0 + 1;
        while (q1 < nq && T + tot[2ull * q1 + what] <= budget) T += tot[2ull * q1 + what], V += tot[2ull * q1 + 1], ++q1;
        if (V) {
            part_q0.push_back(q0);
            part_q1.push_back(q1);
        }
        q0 = q1;
    }
    if (by_slice) return ids_by_slice(c, d_keys, h_qoff, nq, d_ids, h_id_off, counts, tot, part_q0, part_q1,
                                      (uint32_t)S);
    for (size_t pi = 0; pi < part_q0.size(); ++pi) {
        const uint32_t q0 = part_q0[pi], q1 = part_q1[pi];
        uint64_t T = 0, W = 0;
        for (uint32_t q = q0; q < q1; ++q) T += tot[2ull * q + 1], W += tot[2ull * q];
        const uint64_t k0 = h_qoff[q0], Kp = h_qoff[q1] - k0;
        const uint64_t nch = (W + kIdsChunk - 1) / kIdsChunk;
        // record numbering, chunk value counts, chunk bases
        AB_TRY(c->koff.ensure((Kp + 1) * 8));
        AB_TRY(c->parts.ensure((scan_parts_needed(std::max(std::max(Kp, nch), T)) + 2) * 8));
        AB_TRY(c->id_cbase.ensure((nch + 1) * 8));
        AB_TRY(c->id_ccnt.ensure(nch * 8 + 8));
        AB_TRY(c->id_vals.ensure(T * 8 + 64));
        AB_TRY(c->id_sorted.ensure(T * 8 + 64));
        AB_TRY(c->id_start.ensure((size_t)(q1 - q0) * 8 + 8));
        launch_scan_u64((const uint64_t *)c->cnt.p + k0, (uint64_t *)c->koff.p, Kp, (uint64_t *)c->parts.p, s);
        const uint32_t grid = (uint32_t)std::min<uint64_t>(nch, kIdsGrid);
        const uint32_t *keys_p = d_keys + k0, *kq_p = (const uint32_t *)c->kq.p + k0;
        hipLaunchKernelGGL(k_ids_count, dim3(grid), dim3(kIdsNT), 0, s, d, keys_p, kq_p,
                           (const uint64_t *)c->koff.p, Kp, W, (uint64_t *)c->id_ccnt.p);
        HIP_TRY(hipGetLastError());
        launch_scan_u64((const uint64_t *)c->id_ccnt.p, (uint64_t *)c->id_cbase.p, nch, (uint64_t *)c->parts.p, s);
        // packed (query - q0) << vbits | value, sorted query-major with values
        // ascending; u32 when both fit
        const int qbits = bits_for(q1 - q0 - 1);
        if (vbits < 32 && vbits + qbits <= 32)
            AB_TRY(sort_and_write<uint32_t>(c, d, keys_p, kq_p, Kp, W, grid, q0, vbits, vbits + qbits, T, d_ids,
                                            counts));
        else
            AB_TRY(sort_and_write<unsigned long long>(c, d, keys_p, kq_p, Kp, W, grid, q0, vbits, vbits + qbits, T,
                                                      d_ids, counts));
    }
    return AB_OK;
}
