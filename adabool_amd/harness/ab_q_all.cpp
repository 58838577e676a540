// ab_q_all -- every query file of a directory through both methods and the
// dispatched path, checked against the ground truth, with per-method
// prepare and execution statistics.  The C++ counterpart of
// test/adsfilter_q_all.cu:322-524 over adabool_ivf.hpp:
//
//   ab_q_all [filter_file] [query_dir]
//
// filter_file: "key value" lines (ReadCOO); query_dir: query_vec_<n>.coo
// files of one key per line, taken in FindQueryFiles order.  The key and
// value spaces are max + 1 of the pairs, as in the reference.  Exit status 0
// when every method matched the ground truth on every query.
//
// Differences from the reference harness: results are written in full by the
// library, so no memset precedes a timed call; the third method ("auto") is
// the per-query dispatch (the threshold rule) with ids compacted on the
// device, which the reference did not have.
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

struct Row {
    std::string name;
    size_t qsize = 0, rsize = 0;
    double prep[2] = {0, 0}, exec[3] = {0, 0, 0};
    bool ok[3] = {false, false, false};
    int chosen = 0;
};

void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

std::string base_name(const std::string &p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

}  // namespace

int main(int argc, char **argv)
{
    std::string filter_path = "../dataset/adsfilter/filter.coo";
    std::string query_dir = "../dataset/adsfilter";
    if (argc >= 2) filter_path = argv[1];
    if (argc >= 3) query_dir = argv[2];
    std::printf("Filter file: %s\nQuery directory: %s\n", filter_path.c_str(), query_dir.c_str());

    const std::vector<std::string> files = adabool::FindQueryFiles(query_dir);
    if (files.empty()) {
        std::fprintf(stderr, "No query files found in %s\n", query_dir.c_str());
        return 1;
    }
    std::printf("Found %zu query files\n", files.size());

    std::vector<uint32_t> keys, vals;
    uint32_t max_row = 0, max_col = 0;
    auto t0 = Clock::now();
    if (!harness::read_coo(filter_path, keys, vals, max_row, max_col)) return 1;
    auto t1 = Clock::now();
    std::printf("Read filter data: %zu entries, max_row=%u, max_col=%u (%.3f ms)\n", keys.size(), max_row, max_col,
                ms_between(t0, t1));
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
        const size_t res_bytes = (ncols + 7) / 8;                       // the reference's result size
        const size_t stride = ab_result_stride_bytes(list.index());    // the dispatched path's bitmap
        uint8_t *d_res[2];
        uint32_t *d_auto = nullptr, *d_ids = nullptr, *d_cnt = nullptr;
        for (auto &p : d_res) check_hip(hipMalloc((void **)&p, res_bytes), "hipMalloc");
        check_hip(hipMalloc((void **)&d_auto, stride), "hipMalloc");
        check_hip(hipMalloc((void **)&d_ids, (size_t)ncols * 4), "hipMalloc");
        check_hip(hipMalloc((void **)&d_cnt, 4), "hipMalloc");
        std::vector<Row> rows;
        std::vector<double> prep[2], exec[3];
        std::vector<uint8_t> host(std::max(res_bytes, stride));
        for (size_t qi = 0; qi < files.size(); ++qi) {
            Row r;
            r.name = base_name(files[qi]);
            std::printf("\n[%zu/%zu] Processing %s\n", qi + 1, files.size(), r.name.c_str());
            std::vector<uint32_t> q;
            if (!harness::read_query(files[qi], q)) {
                std::fprintf(stderr, "Failed to read query file: %s\n", files[qi].c_str());
                continue;
            }
            r.qsize = q.size();
            const std::vector<uint32_t> want = truth.ids(q);
            r.rsize = want.size();
            // the dense method's key bitmap: keys inside the key space only
            std::vector<uint8_t> qbm((nrows + 7) / 8, 0);
            for (uint32_t k : q)
                if (k < nrows) qbm[k / 8] |= (uint8_t)(1u << (k % 8));

            t0 = Clock::now();
            auto dw = list.Prepare(qbm.size(), qbm.data());
            t1 = Clock::now();
            r.prep[0] = ms_between(t0, t1);
            t0 = Clock::now();
            (void)adabool::ListUnion_dense(*dw, (uint32_t *)d_res[0], stream);
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            t1 = Clock::now();
            r.exec[0] = ms_between(t0, t1);

            t0 = Clock::now();
            auto sw = list.Prepare(q);
            t1 = Clock::now();
            r.prep[1] = ms_between(t0, t1);
            t0 = Clock::now();
            (void)adabool::ListUnion_sparse(*sw, (uint32_t *)d_res[1], stream);
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            t1 = Clock::now();
            r.exec[1] = ms_between(t0, t1);

            t0 = Clock::now();
            r.chosen = adabool::Query(list, q, d_auto, d_ids, d_cnt, stream);
            check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
            t1 = Clock::now();
            r.exec[2] = ms_between(t0, t1);

            for (int m = 0; m < 2; ++m) {
                check_hip(hipMemcpy(host.data(), d_res[m], res_bytes, hipMemcpyDeviceToHost), "hipMemcpy");
                r.ok[m] = harness::collect_ids(host.data(), res_bytes, ncols) == want;
            }
            uint32_t n = 0;
            check_hip(hipMemcpy(&n, d_cnt, 4, hipMemcpyDeviceToHost), "hipMemcpy");
            std::vector<uint32_t> ids(n);
            if (n) check_hip(hipMemcpy(ids.data(), d_ids, (size_t)n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            check_hip(hipMemcpy(host.data(), d_auto, stride, hipMemcpyDeviceToHost), "hipMemcpy");
            r.ok[2] = ids == want && harness::collect_ids(host.data(), stride, ncols) == want;

            for (int m = 0; m < 2; ++m) prep[m].push_back(r.prep[m]);
            for (int m = 0; m < 3; ++m) exec[m].push_back(r.exec[m]);
            std::printf("  Query size: %zu, Result size: %zu\n", r.qsize, r.rsize);
            std::printf("  Dense:  prepare=%.3f ms, exec=%.3f ms, correct=%s\n", r.prep[0], r.exec[0],
                        r.ok[0] ? "PASS" : "FAIL");
            std::printf("  Sparse: prepare=%.3f ms, exec=%.3f ms, correct=%s\n", r.prep[1], r.exec[1],
                        r.ok[1] ? "PASS" : "FAIL");
            std::printf("  Auto:   exec=%.3f ms (%s), correct=%s\n", r.exec[2],
                        r.chosen == AB_PATH_DENSE ? "dense" : "sparse", r.ok[2] ? "PASS" : "FAIL");
            rows.push_back(r);
        }
        check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");

        size_t pass[3] = {0, 0, 0};
        for (const Row &r : rows)
            for (int m = 0; m < 3; ++m) pass[m] += r.ok[m];
        std::printf("\n========================================\n           SUMMARY REPORT\n"
                    "========================================\nTotal queries processed: %zu\n\n",
                    rows.size());
        std::printf("Correctness Summary:\n  Dense method:  %zu/%zu PASS\n  Sparse method: %zu/%zu PASS\n"
                    "  Auto method:   %zu/%zu PASS\n\n",
                    pass[0], rows.size(), pass[1], rows.size(), pass[2], rows.size());
        std::printf("========================================\n");
        harness::print_stats("Dense Method - Prepare Time:", prep[0]);
        harness::print_stats("\nDense Method - Execution Time:", exec[0]);
        std::printf("\n========================================\n");
        harness::print_stats("Sparse Method - Prepare Time:", prep[1]);
        harness::print_stats("\nSparse Method - Execution Time:", exec[1]);
        std::printf("\n========================================\n");
        harness::print_stats("Auto Method (dispatch + ids) - Execution Time:", exec[2]);
        std::printf("\n========================================\nDetailed Results:\n"
                    "----------------------------------------\n");
        std::printf("Query                  | QSize    | RSize    | Dense Prep | Dense Exec | Sparse Prep | "
                    "Sparse Exec | Auto Exec | Path   | D_Correct | S_Correct | A_Correct\n");
        for (const Row &r : rows)
            std::printf("%-22s | %8zu | %8zu | %10.3f | %10.3f | %11.3f | %11.3f | %9.3f | %-6s | %-9s | %-9s | "
                        "%-9s\n",
                        r.name.c_str(), r.qsize, r.rsize, r.prep[0], r.exec[0], r.prep[1], r.exec[1], r.exec[2],
                        r.chosen == AB_PATH_DENSE ? "dense" : "sparse", r.ok[0] ? "PASS" : "FAIL",
                        r.ok[1] ? "PASS" : "FAIL", r.ok[2] ? "PASS" : "FAIL");
        const bool all = pass[0] == rows.size() && pass[1] == rows.size() && pass[2] == rows.size();
        std::printf("\n========================================\nFinal result: %s\n"
                    "========================================\n",
                    all ? "ALL TESTS PASSED" : "SOME TESTS FAILED");
        for (auto &p : d_res) (void)hipFree(p);
        (void)hipFree(d_auto);
        (void)hipFree(d_ids);
        (void)hipFree(d_cnt);
        // the stream outlives the list's context on it: it ends with the process
        return all ? 0 : 1;
    } catch (const adabool::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
