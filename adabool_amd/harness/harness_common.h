// harness_common.h -- pieces shared by the two command-line harnesses: the
// COO and query readers, the ground truth, bitmap -> ids, timing and the
// statistics print-out, all over the C ABI.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "adabool.h"

namespace harness {

using Clock = std::chrono::high_resolution_clock;
inline double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// ReadCOO (test/adsfilter_q_all.cu:127-161) through ab_read_coo.
inline bool read_coo(const std::string &path, std::vector<uint32_t> &keys, std::vector<uint32_t> &vals,
                     uint32_t &max_row, uint32_t &max_col)
{
    const int64_t n = ab_read_coo(path.c_str(), nullptr, nullptr, 0, &max_row, &max_col);
    if (n < 0) {
        std::fprintf(stderr, "Failed to open file: %s\n", path.c_str());
        return false;
    }
    keys.resize((size_t)n);
    vals.resize((size_t)n);
    ab_read_coo(path.c_str(), keys.data(), vals.data(), (uint64_t)n, &max_row, &max_col);
    return true;
}

// ReadQuery (test/adsfilter_q_all.cu:164-193) through ab_read_query.
inline bool read_query(const std::string &path, std::vector<uint32_t> &q)
{
    const int64_t n = ab_read_query(path.c_str(), nullptr, 0);
    if (n < 0) {
        std::fprintf(stderr, "Failed to open file: %s\n", path.c_str());
        return false;
    }
    q.resize((size_t)n);
    ab_read_query(path.c_str(), q.data(), (uint64_t)n);
    return true;
}

// The predicate of ComputeGroundTruth (test/adsfilter_q_all.cu:196-208),
// {v : (k, v) a pair, k in q}, evaluated through a by-key copy of the pairs
// instead of a scan of every pair per query.
class GroundTruth {
public:
    GroundTruth(const std::vector<uint32_t> &keys, const std::vector<uint32_t> &vals, uint32_t nkeys,
                uint32_t nvalues)
        : nkeys_(nkeys), nvalues_(nvalues), off_((size_t)nkeys + 1, 0), vals_(keys.size()), mark_(nvalues, 0)
    {
        for (uint32_t k : keys) off_[(size_t)k + 1]++;
        for (size_t k = 0; k < nkeys; ++k) off_[k + 1] += off_[k];
        std::vector<uint64_t> cur(off_.begin(), off_.end() - 1);
        for (size_t i = 0; i < keys.size(); ++i) vals_[cur[keys[i]]++] = vals[i];
    }
    // ascending distinct values of q's keys (keys outside the key space add nothing)
    std::vector<uint32_t> ids(const std::vector<uint32_t> &q)
    {
        std::vector<uint32_t> out;
        for (uint32_t k : q) {
            if (k >= nkeys_) continue;
            for (uint64_t j = off_[k]; j < off_[(size_t)k + 1]; ++j) {
                const uint32_t v = vals_[j];
                if (!mark_[v]) {
                    mark_[v] = 1;
                    out.push_back(v);
                }
            }
        }
        for (uint32_t v : out) mark_[v] = 0;
        std::sort(out.begin(), out.end());
        return out;
    }

private:
    uint32_t nkeys_, nvalues_;
    std::vector<uint64_t> off_;
    std::vector<uint32_t> vals_;
    std::vector<uint8_t> mark_;
};

// CollectIdsFromBitmap (test/adsfilter_batch.cu:82-97): LSB-first bytes,
// ids >= ncols dropped.
inline std::vector<uint32_t> collect_ids(const uint8_t *bm, size_t bytes, uint32_t ncols)
{
    std::vector<uint32_t> out;
    for (size_t b = 0; b < bytes; ++b)
        for (uint32_t x = bm[b]; x; x &= x - 1) {
            const uint32_t id = (uint32_t)(b * 8 + __builtin_ctz(x));
            if (id < ncols) out.push_back(id);
        }
    return out;
}

// CalculateStats + PrintStats (test/adsfilter_q_all.cu:258-321) through ab_stats.
inline void print_stats(const char *title, std::vector<double> v)
{
    double s[9];
    ab_stats(v.data(), v.size(), s);
    std::printf("%s\n", title);
    const char *names[9] = {"Mean:  ", "Median:", "Stddev:", "Min:   ", "Max:   ", "P50:   ", "P90:   ",
                            "P95:   ", "P99:   "};
    for (int i = 0; i < 9; ++i) std::printf("  %s %.3f ms\n", names[i], s[i]);
}

}  // namespace harness
