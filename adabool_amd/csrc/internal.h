// internal.h -- shared declarations of the AdaBool MI355X engine (host side).
//
// Two device layouts of the same compressed lists are kept resident:
//   key-major   (KM): the reference's per-type arrays ordered by (key, window)
//                     with key_segments prefix counts; used by the sparse path.
//   value-tiled (VT): the same records re-ordered by (value tile, key, window)
//                     with a delta-coded key stream in 1024-record blocks; used
//                     by the dense path, whose work-groups each own one value
//                     tile and write it exactly once.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "adabool.h"
#include "layout.h"

namespace ab {

constexpr int kTypes = AB_NTYPES;
constexpr int kPayloadBytes[kTypes] = {0, 4, 8, 16, 32, 32};
// kMinNum of ArraySizeTrait (src/AdaBool/inverted_list.cu.h:21-52): payload
// bytes emitted unconditionally before the zero terminator applies.
constexpr int kMinNum[kTypes] = {0, 1, 5, 9, 17, 0};


// Block descriptor of the value-tiled key stream.
struct BlockMeta {
    uint32_t start_key;  // key of the block's first record
    uint32_t off_kb;     // offset of the block's 1024*w delta bytes, in units of
                         // 1024 bytes (every block is a multiple of that long)
    uint32_t tw;         // type | log2(width) << 4 | nvalid << 8
    uint32_t rec_base;   // first record (index into the type's VT arrays)
};

struct HostIndex {
    uint32_t nkeys = 0, nvalues = 0;
    uint64_t nnz = 0;
    std::vector<uint32_t> seg8;  // (nkeys+1) * 8
    // key-major
    std::vector<uint32_t> hdr[kTypes];
    std::vector<uint8_t> pay[kTypes];
    // value-tiled
    uint32_t tile_shift = 12, ntiles = 0;
    std::vector<uint32_t> vhdr[kTypes];
    std::vector<uint8_t> vpay[kTypes];
    std::vector<BlockMeta> vblk;
    std::vector<uint8_t> vkeys;
    std::vector<uint32_t> tile_blk;  // ntiles+1
    uint64_t nrec(int t) const { return hdr[t].size(); }
    uint64_t nrec_total() const {
        uint64_t s = 0;
        for (int t = 0; t < kTypes; ++t) s += hdr[t].size();
        return s;
    }
    uint64_t list_bytes() const {
        uint64_t s = 0;
        for (int t = 0; t < kTypes; ++t) s += hdr[t].size() * (4 + kPayloadBytes[t]);
        return s;
    }
};

// Error reporting (thread-local message).
void set_error(const std::string &msg);

// Host index build (host_index.cpp).  Returns AB_OK or an error status.
ab_status build_host_index(const uint32_t *keys, const uint32_t *vals, uint64_t nnz,
                           uint32_t nkeys, uint32_t nvalues, uint32_t tile_shift,
                           int nthreads, HostIndex *out);
uint32_t auto_tile_shift(uint32_t nvalues);
// Host features of one query (host_index.cpp); same definition as the device
// feature kernel.
void host_features(const uint32_t *seg8, uint32_t nkeys, uint32_t nvalues, uint64_t R,
                   const uint32_t *keys, uint32_t n, float *f, uint64_t *m);

// Every invariant the kernels rely on, checked over a loaded index in one
// pass (index files from outside the process); false with a reason if any
// fails.
bool validate_index(const HostIndex &h, std::string &why);

uint64_t stride_bytes(uint32_t nvalues);
int default_threads();
// COO text file -> pairs in file order (ReadCOO rules), parsed in parallel.
bool read_coo_file(const char *path, std::vector<uint32_t> &keys, std::vector<uint32_t> &vals, uint32_t &max_key,
                   uint32_t &max_value);

// Parallel for over [0, n) split into contiguous ranges.
template <class F>
void parallel_ranges(uint64_t n, int nthreads, F &&f);

}  // namespace ab

// Host-side opaque handles.
struct ab_dataset {
    std::vector<uint32_t> a, b;
    std::vector<uint64_t> off;
    uint64_t count = 0;
};

#include <thread>
namespace ab {
template <class F>
void parallel_ranges(uint64_t n, int nthreads, F &&f)
{
    if (nthreads <= 1 || n < 2) {
        f(0, n, 0);
        return;
    }
    if ((uint64_t)nthreads > n) nthreads = (int)n;
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        uint64_t b = n * t / nthreads, e = n * (t + 1) / nthreads;
        th.emplace_back([&f, b, e, t]() { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}
}  // namespace ab
