// ab_batch -- the query files of a directory in fixed-size batches through
// batched dense, batched sparse and the dispatched batch path, each batch
// checked against the ground truth, with per-method throughput.  The C++
// counterpart of test/adsfilter_batch.cu:305-599 over adabool_ivf.hpp:
//
//   ab_batch [filter_file] [query_dir] [batch_size]   (batch_size > 0, default 4)
//
// As in the reference, only complete batches run (the file list is cut to a
// multiple of batch_size) and throughput is queries * 1000 / total execution
// ms.  Differences: the filter file is read (the reference's ReadCOO in this
// harness is a stub that returns no pairs, test/adsfilter_batch.cu:109-113);
// any batch size works (the reference's dense batch handled 1..32); the third
// method ("auto") is one device-dispatched call per batch (threshold rule,
// fused union launch, ids compacted on the device).  Exit status 0 when
// every batch of every method matched.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "adabool_ivf.hpp"
#include "harness_common.h"

using harness::Clock;
using harness::ms_between;

namespace {

void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

uint64_t align16(uint64_t x) { return (x + 15) & ~15ull; }

void usage(const char *prog)
{
    std::printf("Usage: %s [filter_file] [query_dir] [batch_size]\n"
                "  filter_file: Path to filter.coo file (default: ../dataset/adsfilter/filter.coo)\n"
                "  query_dir:   Directory containing query files (default: ../dataset/adsfilter)\n"
                "  batch_size:  Number of queries per batch (default: 4, must be > 0)\n",
                prog);
}

}  // namespace

int main(int argc, char **argv)
{
    std::string filter_path = "../dataset/adsfilter/filter.coo";
    std::string query_dir = "../dataset/adsfilter";
    uint32_t batch = 4;
    if (argc >= 2) filter_path = argv[1];
    if (argc >= 3) query_dir = argv[2];
    if (argc >= 4) {
        batch = (uint32_t)std::strtoul(argv[3], nullptr, 10);
        if (batch == 0) {
            std::fprintf(stderr, "Error: batch_size must be greater than 0\n");
            usage(argv[0]);
            return 1;
        }
    }
    std::printf("========================================\n  AdsFilter Batch Query Test\n"
                "========================================\nFilter file: %s\nQuery directory: %s\nBatch size: %u\n",
                filter_path.c_str(), query_dir.c_str(), batch);

    std::vector<std::string> files = adabool::FindQueryFiles(query_dir);
    if (files.empty()) {
        std::fprintf(stderr, "No query files found in %s\n", query_dir.c_str());
        return 1;
    }
    std::printf("Found %zu query files\n", files.size());
    const size_t complete = files.size() / batch * batch;
    if (complete < files.size()) {
        std::printf("WARNING: Truncating to %zu queries (complete batches only)\n", complete);
        files.resize(complete);
    }
    if (files.empty()) {
        std::fprintf(stderr, "Not enough queries to form a complete batch (need at least %u queries)\n", batch);
        return 1;
    }

    std::vector<uint32_t> keys, vals;
    uint32_t max_row = 0, max_col = 0;
    auto t0 = Clock::now();
    if (!harness::read_coo(filter_path, keys, vals, max_row, max_col)) return 1;
    auto t1 = Clock::now();
    std::printf("Read filter data: %zu entries, rows=%u, cols=%u (%.3f ms)\n", keys.size(), max_row + 1,
                max_col + 1, ms_between(t0, t1));
    const uint32_t nrows = max_row + 1, ncols = max_col + 1;

    try {
        t0 = Clock::now();
        adabool::ivf_list list;
        list.BuildFrom(keys.data(), vals.data(), keys.size(), nrows, ncols);
        t1 = Clock::now();
        std::printf("Built inverted index (%.3f ms)\n", ms_between(t0, t1));
        harness::GroundTruth truth(keys, vals, nrows, ncols);

        hipStream_t stream;
        check_hip(hipStreamCreate(&stream), "hipStreamCreate");
        const uint64_t res_aligned = align16((ncols + 7) / 8);  // the reference's per-query result stride
        const uint64_t qbytes = align16((nrows + 7) / 8);
        const uint64_t stride = ab_result_stride_bytes(list.index());
        uint8_t *d_dense = nullptr, *d_sparse = nullptr, *d_bm = nullptr;
        uint32_t *d_ids = nullptr, *d_cnt = nullptr, *d_qidx = nullptr, *d_keys = nullptr;
        check_hip(hipMalloc((void **)&d_dense, batch * res_aligned), "hipMalloc");
        check_hip(hipMalloc((void **)&d_sparse, batch * res_aligned), "hipMalloc");
        check_hip(hipMalloc((void **)&d_bm, batch * stride), "hipMalloc");
        check_hip(hipMalloc((void **)&d_ids, (uint64_t)batch * ncols * 4), "hipMalloc");
        check_hip(hipMalloc((void **)&d_cnt, (uint64_t)batch * 4), "hipMalloc");

        std::printf("\nPerforming warmup...\n");
        {
            std::vector<uint8_t> zero(batch * qbytes, 0);
            auto w = list.PrepareBatch(batch, qbytes, zero.data());
            w->RunArrays(stream, (uint32_t *)d_dense, (uint32_t)(res_aligned / 4));
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            std::printf("Warmup completed.\n");
        }

        const size_t nb = files.size() / batch;
        size_t pass[3] = {0, 0, 0};
        double total[3] = {0, 0, 0};
        std::vector<uint8_t> host(batch * std::max(res_aligned, stride));
        std::vector<uint32_t> host_cnt(batch);
        for (size_t b = 0; b < nb; ++b) {
            const size_t s0 = b * batch;
            std::printf("\n========================================\nProcessing Batch %zu/%zu (queries %zu-%zu)\n",
                        b + 1, nb, s0, s0 + batch - 1);
            std::vector<std::vector<uint32_t>> qs(batch);
            for (uint32_t i = 0; i < batch; ++i) {
                if (!harness::read_query(files[s0 + i], qs[i])) {
                    std::fprintf(stderr, "Failed to read query: %s\n", files[s0 + i].c_str());
                    return 1;
                }
                std::printf("  Query %zu: %zu keys\n", s0 + i, qs[i].size());
            }
            std::vector<std::vector<uint32_t>> want(batch);
            for (uint32_t i = 0; i < batch; ++i) want[i] = truth.ids(qs[i]);

            // dense: query-major key bitmaps, 16-byte rows
            std::vector<uint8_t> qbm(batch * qbytes, 0);
            for (uint32_t i = 0; i < batch; ++i)
                for (uint32_t k : qs[i])
                    if (k < nrows) qbm[i * qbytes + k / 8] |= (uint8_t)(1u << (k % 8));
            t0 = Clock::now();
            auto dw = list.PrepareBatch(batch, qbytes, qbm.data());
            t1 = Clock::now();
            const double prep_dense = ms_between(t0, t1);
            // sparse: flattened keys, the query of each key on the device
            std::vector<uint32_t> flat, qidx;
            for (uint32_t i = 0; i < batch; ++i)
                for (uint32_t k : qs[i]) {
                    flat.push_back(k);
                    qidx.push_back(i);
                }
            t0 = Clock::now();
            auto sw = list.Prepare(flat);
            if (d_qidx) (void)hipFree(d_qidx);
            check_hip(hipMalloc((void **)&d_qidx, flat.size() * 4 + 4), "hipMalloc");
            if (!flat.empty())
                check_hip(hipMemcpy(d_qidx, qidx.data(), qidx.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
            t1 = Clock::now();
            const double prep_sparse = ms_between(t0, t1);
            // the dispatched batch: keys on the device, host offsets
            std::vector<uint64_t> qoff(batch + 1, 0);
            for (uint32_t i = 0; i < batch; ++i) qoff[i + 1] = qoff[i] + qs[i].size();
            if (d_keys) (void)hipFree(d_keys);
            check_hip(hipMalloc((void **)&d_keys, flat.size() * 4 + 4), "hipMalloc");
            if (!flat.empty())
                check_hip(hipMemcpy(d_keys, flat.data(), flat.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");

            double exec[3];
            t0 = Clock::now();
            dw->RunArrays(stream, (uint32_t *)d_dense, (uint32_t)(res_aligned / 4));
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            t1 = Clock::now();
            exec[0] = ms_between(t0, t1);
            t0 = Clock::now();
            (void)adabool::BatchListUnion_sparse(*sw, (uint32_t *)d_sparse, d_qidx, (uint32_t)res_aligned, stream);
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            t1 = Clock::now();
            exec[1] = ms_between(t0, t1);
            t0 = Clock::now();
            adabool::check(ab_query_batch(list.ctx(stream), d_keys, qoff.data(), batch, AB_PATH_AUTO,
                                          (uint32_t *)d_bm, d_ids, ncols, d_cnt, nullptr),
                           "ab_query_batch");
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            t1 = Clock::now();
            exec[2] = ms_between(t0, t1);

            bool ok[3] = {true, true, true};
            const char *names[2] = {"Dense", "Sparse"};
            uint8_t *outs[2] = {d_dense, d_sparse};
            for (int m = 0; m < 2; ++m) {
                check_hip(hipMemcpy(host.data(), outs[m], batch * res_aligned, hipMemcpyDeviceToHost), "hipMemcpy");
                for (uint32_t i = 0; i < batch; ++i) {
                    const auto got = harness::collect_ids(host.data() + i * res_aligned, res_aligned, ncols);
                    if (got != want[i]) {
                        ok[m] = false;
                        std::printf("  Query %zu %s MISMATCH: expected %zu, got %zu\n", s0 + i, names[m],
                                    want[i].size(), got.size());
                    }
                }
            }
            check_hip(hipMemcpy(host_cnt.data(), d_cnt, (uint64_t)batch * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            check_hip(hipMemcpy(host.data(), d_bm, batch * stride, hipMemcpyDeviceToHost), "hipMemcpy");
            for (uint32_t i = 0; i < batch; ++i) {
                std::vector<uint32_t> ids(host_cnt[i]);
                if (host_cnt[i])
                    check_hip(hipMemcpy(ids.data(), d_ids + (uint64_t)i * ncols, (uint64_t)host_cnt[i] * 4,
                                        hipMemcpyDeviceToHost),
                              "hipMemcpy");
                if (ids != want[i] || harness::collect_ids(host.data() + i * stride, stride, ncols) != want[i]) {
                    ok[2] = false;
                    std::printf("  Query %zu Auto MISMATCH: expected %zu, got %u\n", s0 + i, want[i].size(),
                                host_cnt[i]);
                }
            }
            for (int m = 0; m < 3; ++m) {
                pass[m] += ok[m];
                total[m] += exec[m];
            }
            std::printf("Batch %zu Results:\n  Dense:  Correct=%s (prepare %.3f ms, exec %.3f ms)\n"
                        "  Sparse: Correct=%s (prepare %.3f ms, exec %.3f ms)\n  Auto:   Correct=%s (exec %.3f ms)\n",
                        b + 1, ok[0] ? "YES" : "NO", prep_dense, exec[0], ok[1] ? "YES" : "NO", prep_sparse,
                        exec[1], ok[2] ? "YES" : "NO", exec[2]);
        }

        std::printf("\n========================================\n           SUMMARY REPORT\n"
                    "========================================\nBatch size: %u\nTotal batches processed: %zu\n"
                    "Total queries processed: %zu\n\n",
                    batch, nb, files.size());
        std::printf("Correctness Summary:\n  Dense method:  %zu/%zu batches PASS\n  Sparse method: %zu/%zu batches "
                    "PASS\n  Auto method:   %zu/%zu batches PASS\n\n",
                    pass[0], nb, pass[1], nb, pass[2], nb);
        std::printf("========================================\nThroughput Results:\n"
                    "----------------------------------------\n");
        const char *label[3] = {"Dense method: ", "Sparse method:", "Auto method:  "};
        for (int m = 0; m < 3; ++m)
            std::printf("  %s %.2f queries/sec (total exec time: %.3f ms)\n", label[m],
                        total[m] > 0 ? files.size() * 1000.0 / total[m] : 0.0, total[m]);
        const bool all = pass[0] == nb && pass[1] == nb && pass[2] == nb;
        std::printf("\n========================================\nFinal result: %s\n"
                    "========================================\n",
                    all ? "ALL TESTS PASSED" : "SOME TESTS FAILED");
        for (void *p : {(void *)d_dense, (void *)d_sparse, (void *)d_bm, (void *)d_ids, (void *)d_cnt,
                        (void *)d_qidx, (void *)d_keys})
            if (p) (void)hipFree(p);
        return all ? 0 : 1;
    } catch (const adabool::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
