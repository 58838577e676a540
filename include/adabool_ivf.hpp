/*
 * adabool_ivf.hpp -- the reference's C++ calling convention over the C ABI.
 *
 * The reference is a header-only template library: callers build an
 * ivf_list, Prepare a worker per query (keys for the sparse method, a key
 * bitmap for the dense one) and run the per-type launches through free
 * functions on a stream of their own (src/AdaBool/inverted_list.h:315-605,
 * test/adsfilter_q_all.cu:26-105, test/adsfilter_batch.cu:28-79).  This
 * header keeps those names and argument meanings so that such a caller can
 * switch by changing its includes:
 *
 *   reference                                   here
 *   ivf_list<256,64>::BuildFrom(kvs)            adabool::ivf_list::BuildFrom(kvs)
 *   Prepare(keys) -> InvertedListWorker         same (the keys are kept on the host)
 *   Prepare(len, bitmap)                        same (the bitmap is kept on the host)
 *   PrepareBatch / PrepareBatchDevice           same -> InvertedListBatchWorkerImpl
 *   ListUnion_sparse(worker, d_res, stream)     same, ab_list_union_sparse
 *   ListUnion_dense(worker, d_res, stream)      same, ab_list_union_dense
 *   worker.RunArrays(stream, d_res, len_u32)    same, ab_list_union_dense_batch
 *   BatchListUnion_sparse(w, d_res, d_qidx,     same, ab_list_union_sparse_batch
 *                         res_len_bytes, stream)
 *
 * Result buffers keep the reference's sizes: a single query's d_res holds
 * ceil(ncols / 8) bytes (test/adsfilter_q_all.cu:361-368) and exactly that
 * many are written; batched results are res_len apart.  Results are written
 * in full, so the callers' memsets are no longer needed (they are harmless).
 * Errors: the reference aborts on CUDA errors; here every failure throws
 * adabool::Error (status and message of ab_last_error()), which terminates a
 * caller that does not catch it.
 *
 * One ab_ctx is made per stream a caller passes (on first use) and kept by
 * the ivf_list.  Stream 0, the reference's default, gets a context stream of
 * its own, and every call on it returns only when its work is done, so that
 * the caller's later default-stream work (a hipMemcpy of the result) finds
 * it finished; workers borrow the ivf_list, which must outlive them, as in
 * the reference (src/AdaBool/inverted_list.h:584).  Not thread-safe across
 * host threads, as the reference is not.
 */
#ifndef ADABOOL_IVF_HPP
#define ADABOOL_IVF_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "adabool.h"

namespace adabool {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &what) : std::runtime_error(what), status(status) {}
    int status;
};

inline void check(ab_status s, const char *call)
{
    if (s != AB_OK) throw Error(s, std::string(call) + ": " + ab_status_str(s) + ": " + ab_last_error());
}

inline void check_hip(hipError_t e, const char *call)
{
    if (e != hipSuccess) throw Error(AB_DEVICE_ERROR, std::string(call) + ": " + hipGetErrorString(e));
}

class ivf_list;

// Prepare(keys) (src/AdaBool/inverted_list.h:481-527): the query's keys,
// duplicates and keys >= nkeys included; extents and chunking run on the
// device when the union is launched.
class InvertedListWorker {
public:
    InvertedListWorker(const ivf_list &l, std::vector<uint32_t> keys) : list(&l), keys(std::move(keys)) {}
    const ivf_list *list;
    std::vector<uint32_t> keys;
};

// Prepare(len, bitmap) (src/AdaBool/inverted_list.h:530-543): the query's key
// bitmap, LSB-first bytes, bit k = key k.
class InvertedListBitmapWorkerImpl {
public:
    InvertedListBitmapWorkerImpl(const ivf_list &l, const uint8_t *bm, uint64_t len)
        : list(&l), bitmap(bm, bm + len) {}
    const ivf_list *list;
    std::vector<uint8_t> bitmap;
};

// PrepareBatch / PrepareBatchDevice (src/AdaBool/inverted_list.h:564-605):
// batch_size device bitmaps of len bytes each, query-major.
class InvertedListBatchWorkerImpl {
public:
    InvertedListBatchWorkerImpl(const ivf_list &l, uint64_t batch, uint64_t len, const uint8_t *d_bitmaps,
                                void *owned)
        : list(&l), batch_size(batch), len(len), d_bitmaps(d_bitmaps), owned_(owned) {}
    ~InvertedListBatchWorkerImpl()
    {
        if (owned_) (void)hipFree(owned_);
    }
    InvertedListBatchWorkerImpl(const InvertedListBatchWorkerImpl &) = delete;
    InvertedListBatchWorkerImpl &operator=(const InvertedListBatchWorkerImpl &) = delete;
    // RunArrays (src/AdaBool/inverted_list.cu.h:224-228): query i's result at
    // d_res + i * res_len_u32, for any batch size.
    inline void RunArrays(hipStream_t stream, uint32_t *d_res, uint32_t res_len_u32) const;

    const ivf_list *list;
    uint64_t batch_size, len;
    const uint8_t *d_bitmaps;

private:
    void *owned_;
};

class ivf_list {
public:
    ivf_list() = default;
    ivf_list(const ivf_list &) = delete;
    ivf_list &operator=(const ivf_list &) = delete;
    ~ivf_list()
    {
        for (auto &kv : ctx_) ab_ctx_free(kv.second);
        if (ix_) ab_index_free(ix_);
    }

    // BuildFrom (src/AdaBool/inverted_list.h:348-374): any iterable of
    // (key, value) pairs; key and value spaces are max + 1, as the
    // reference's callers derive them (test/adsfilter_q_all.cu:359-360).
    template <class Container> void BuildFrom(const Container &kvs, uint32_t tile_shift = 0)
    {
        std::vector<uint32_t> k, v;
        for (const auto &p : kvs) {
            k.push_back(p.first);
            v.push_back(p.second);
        }
        BuildFrom(k.data(), v.data(), k.size(), 0, 0, tile_shift);
    }
    void BuildFrom(const uint32_t *keys, const uint32_t *values, uint64_t n, uint32_t nkeys, uint32_t nvalues,
                   uint32_t tile_shift = 0)
    {
        release();
        check(ab_index_build(keys, values, n, nkeys, nvalues, tile_shift, &ix_), "ab_index_build");
        check(ab_index_get_info(ix_, &info_), "ab_index_get_info");
    }

    template <class Container> std::unique_ptr<InvertedListWorker> Prepare(const Container &keys) const
    {
        return std::unique_ptr<InvertedListWorker>(
            new InvertedListWorker(*this, std::vector<uint32_t>(keys.begin(), keys.end())));
    }
    std::unique_ptr<InvertedListBitmapWorkerImpl> Prepare(uint64_t len_bytes, const uint8_t *host_bitmap) const
    {
        return std::unique_ptr<InvertedListBitmapWorkerImpl>(
            new InvertedListBitmapWorkerImpl(*this, host_bitmap, len_bytes));
    }
    std::unique_ptr<InvertedListBatchWorkerImpl> PrepareBatch(uint64_t batch, uint64_t len_bytes,
                                                              const uint8_t *host_bitmaps) const
    {
        void *d = nullptr;
        const uint64_t bytes = batch * len_bytes;
        check_hip(hipMalloc(&d, bytes ? bytes : 16), "hipMalloc");
        const hipError_t e = hipMemcpy(d, host_bitmaps, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            check_hip(e, "hipMemcpy");
        }
        return std::unique_ptr<InvertedListBatchWorkerImpl>(
            new InvertedListBatchWorkerImpl(*this, batch, len_bytes, (const uint8_t *)d, d));
    }
    std::unique_ptr<InvertedListBatchWorkerImpl> PrepareBatchDevice(uint64_t batch, uint64_t len_bytes,
                                                                    const uint8_t *dev_bitmaps) const
    {
        return std::unique_ptr<InvertedListBatchWorkerImpl>(
            new InvertedListBatchWorkerImpl(*this, batch, len_bytes, dev_bitmaps, nullptr));
    }

    // The context bound to `stream` (made on first use).
    ab_ctx *ctx(hipStream_t stream) const
    {
        if (!ix_) throw Error(AB_INVALID_ARG, "ivf_list: BuildFrom has not run");
        auto it = ctx_.find(stream);
        if (it != ctx_.end()) return it->second;
        ab_ctx *c = nullptr;
        check(ab_ctx_create(ix_, (void *)stream, &c), "ab_ctx_create");
        ctx_[stream] = c;
        return c;
    }
    ab_index *index() const { return ix_; }
    const ab_index_info &info() const { return info_; }
    uint32_t nkeys() const { return info_.nkeys; }
    uint32_t nvalues() const { return info_.nvalues; }
    // the reference's single-query result size, ceil(ncols / 8) bytes
    uint64_t result_bytes() const { return ((uint64_t)info_.nvalues + 7) / 8; }

private:
    void release()
    {
        for (auto &kv : ctx_) ab_ctx_free(kv.second);
        ctx_.clear();
        if (ix_) ab_index_free(ix_);
        ix_ = nullptr;
    }
    ab_index *ix_ = nullptr;
    ab_index_info info_{};
    mutable std::map<hipStream_t, ab_ctx *> ctx_;
};

// A call on stream 0 completes before it returns (see above).
inline void finish_default(const ivf_list &l, hipStream_t stream)
{
    if (!stream) check(ab_ctx_sync(l.ctx(stream)), "ab_ctx_sync");
}

inline void InvertedListBatchWorkerImpl::RunArrays(hipStream_t stream, uint32_t *d_res, uint32_t res_len_u32) const
{
    check(ab_list_union_dense_batch(list->ctx(stream), (uint32_t)batch_size, len, d_bitmaps, d_res, res_len_u32),
          "ab_list_union_dense_batch");
    finish_default(*list, stream);
}

// ListUnion_dense (test/adsfilter_q_all.cu:26-61).
inline hipError_t ListUnion_dense(const InvertedListBitmapWorkerImpl &w, uint32_t *d_res, hipStream_t stream = 0)
{
    check(ab_list_union_dense(w.list->ctx(stream), w.bitmap.size(), w.bitmap.data(), d_res, w.list->result_bytes()),
          "ab_list_union_dense");
    finish_default(*w.list, stream);
    return hipSuccess;
}

// ListUnion_sparse (test/adsfilter_q_all.cu:64-105).
inline hipError_t ListUnion_sparse(const InvertedListWorker &w, uint32_t *d_res, hipStream_t stream = 0)
{
    check(ab_list_union_sparse(w.list->ctx(stream), w.keys.data(), (uint32_t)w.keys.size(), d_res,
                               w.list->result_bytes()),
          "ab_list_union_sparse");
    finish_default(*w.list, stream);
    return hipSuccess;
}

// BatchListUnion_sparse (test/adsfilter_batch.cu:39-79): key i of the
// worker belongs to query d_qidx[i] (device memory, as the reference's
// LenBuffer upload); query q's result at d_res + q * res_len_bytes / 4.
inline hipError_t BatchListUnion_sparse(const InvertedListWorker &w, uint32_t *d_res, const uint32_t *d_qidx,
                                        uint32_t res_len_per_batch_bytes, hipStream_t stream = 0)
{
    const size_t n = w.keys.size();
    std::vector<uint32_t> qidx(n);
    if (n) check_hip(hipMemcpy(qidx.data(), d_qidx, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    const uint32_t nq = n ? *std::max_element(qidx.begin(), qidx.end()) + 1 : 1;
    check(ab_list_union_sparse_batch(w.list->ctx(stream), w.keys.data(), qidx.data(), (uint32_t)n, nq, d_res,
                                     res_len_per_batch_bytes),
          "ab_list_union_sparse_batch");
    finish_default(*w.list, stream);
    return hipSuccess;
}

// What the reference's callers did by hand, as one call: the method chosen
// per query (threshold rule or learned cost model), bitmap and ascending ids.
inline int Query(const ivf_list &l, const std::vector<uint32_t> &keys, uint32_t *d_bitmap, uint32_t *d_ids,
                 uint32_t *d_count, hipStream_t stream = 0, int path = AB_PATH_AUTO)
{
    int chosen = 0;
    check(ab_query(l.ctx(stream), keys.data(), (uint32_t)keys.size(), path, d_bitmap, d_ids, d_count, &chosen),
          "ab_query");
    finish_default(l, stream);
    return chosen;
}

// FindQueryFiles (test/adsfilter_q_all.cu:211-231).
inline std::vector<std::string> FindQueryFiles(const std::string &dir)
{
    uint64_t need = 0;
    const int64_t n = ab_find_query_files(dir.c_str(), nullptr, 0, &need);
    if (n < 0) {
        std::fprintf(stderr, "Error reading directory: %s\n", ab_last_error());
        return {};
    }
    std::vector<char> buf(need + 1);
    ab_find_query_files(dir.c_str(), buf.data(), buf.size(), &need);
    std::vector<std::string> out;
    for (size_t p = 0; p < need && (int64_t)out.size() < n; p += out.back().size() + 1) out.emplace_back(&buf[p]);
    return out;
}

}  // namespace adabool

#endif  // ADABOOL_IVF_HPP
