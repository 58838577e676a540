// kernels.h -- launch wrappers of the AdaBool CDNA4 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace ab {

// Device view of a resident index (both layouts), passed to kernels by value.
struct DevIndex {
    const uint32_t *seg8;      // (nkeys+1) * 8 prefix rows
    const uint32_t *hdr[6];    // key-major headers per type
    const uint8_t *pay[6];     // key-major payloads per type
    const uint32_t *vhdr[6];   // value-tiled headers per type
    const uint8_t *vpay[6];    // value-tiled payloads per type
    const uint4 *vblk;         // BlockMeta per key-stream block
    const uint8_t *vkeys;      // transposed delta bytes
    const uint32_t *tile_blk;  // ntiles+1 block prefix
    uint32_t nkeys, nvalues, tile_shift, ntiles;
};

constexpr int kMaxGroup = 64;
struct DenseGroup {
    uint32_t B;             // queries in the group (1..64)
    uint32_t q[kMaxGroup];  // output slot (query index) of each group member
};

// mode of the dense kernel's key test
enum DenseMode { kModeBitmap = 0, kModeMask32 = 1, kModeMask64 = 2, kModeMask128 = 3, kModeMask256 = 4 };
// query slots per dense group on the device-dispatched path (groups of up to
// 128 queries use two 64-bit mask words)
constexpr uint32_t kGroupSlots = 256;

// scan scratch: parts needed for n items
uint64_t scan_parts_needed(uint64_t n);

void launch_fill_zero(void *p, uint64_t bytes, hipStream_t s);
void launch_zero_qlist(uint32_t *bm, uint64_t stride_words, uint64_t words, const uint32_t *qlist,
                       uint32_t n, hipStream_t s);
void launch_sparse_plan(const DevIndex &ix, const uint32_t *keys, const uint64_t *qoff,
                        const uint8_t *sel, uint32_t nq, uint32_t *kq, uint64_t *cnt, hipStream_t s);
// n_dev set: n is only the capacity (grid size) and the item count is read
// from *n_dev on the device (at most n).
void launch_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *parts, hipStream_t s,
                     const uint64_t *n_dev = nullptr);
void launch_sparse_union(const DevIndex &ix, const uint32_t *keys, const uint32_t *kq,
                         const uint64_t *koff, uint64_t K, uint32_t *out, uint64_t out_stride,
                         uint32_t chunk, uint32_t grid, hipStream_t s);
// For every key k of query g.q[b]: sets bit k of anybm (1 bit per key) and,
// in the mask modes, bit b of mask[k]; clear=1 writes zeros at the same
// places, returning both buffers to all-zero.
void launch_mask_keys(int mode, const uint32_t *keys, const uint64_t *qoff, const DenseGroup &g,
                      void *mask, uint32_t *anybm, uint32_t nkeys, int clear, hipStream_t s);
// Writes mask[k] (u64) and anybm for every k < nkeys from B query bitmaps.
void launch_mask_from_bitmaps(const uint32_t *bitmaps, uint64_t len_u32, uint32_t b0, uint32_t B,
                              void *mask, uint32_t *anybm, uint32_t nkeys, hipStream_t s);
// Dynamic LDS bytes of the dense kernel for a group of B queries in a mode,
// with or without the per-value column masks (group modes only).
uint32_t dense_lds_bytes(int mode, uint32_t B, uint32_t tile_shift, uint32_t threads, bool use_col);
hipError_t launch_dense_tiles(int mode, uint32_t threads, const DevIndex &ix, const void *mask,
                              const uint32_t *anybm, const DenseGroup &g, uint32_t *out,
                              uint64_t out_stride, uint64_t out_words, bool use_col, hipStream_t s);

// Device-dispatched batch: sparse zero-fill, group masks.
// Zeroes the bitmaps of the queries with sel[q] == match.
void launch_zero_sparse(uint32_t *bm, uint64_t stride, uint64_t words, const uint8_t *sel, uint32_t nq,
                        uint8_t match, hipStream_t s);
// Group masks are indexed by the key's rank among the group's keys.  Per
// group, any_stride u32: the rank words (a uint2 per 32 keys: the any-query
// bits, the rank of the first key), then one u32 total per run of 64 rank
// words (used while the ranks are formed).  A batch runs kMaskAny,
// launch_mask_rank, kMaskSet, the union, kMaskClear, and then zeroes the
// rank words.
constexpr int kMaskAny = 0, kMaskClear = 1, kMaskSet = 2;
inline uint64_t mask_rank_words(uint32_t nkeys)
{
    const uint64_t nw = ((uint64_t)nkeys + 31) / 32;
    return (nw + 1) * 2 + ((nw + 63) / 64 + 3) / 4 * 4;
}
// kq: the query of each key (k_sparse_plan); Kd: the key count on the device
// (graph replays) or NULL for K.
void launch_mask_groups(int mode, const uint32_t *keys, const uint32_t *kq, const uint64_t *Kd, uint64_t K,
                        const uint32_t *qgroup, void *masks, uint64_t mask_stride, uint32_t *anybms,
                        uint64_t any_stride, uint32_t nkeys, int op, hipStream_t s);
void launch_mask_rank(uint32_t *anybms, uint64_t any_stride, uint32_t nkeys, uint32_t ngroups, const uint32_t *plan,
                      hipStream_t s);
// ---- write-once sparse work of device-dispatched batches ----------------
// A sparse query's bitmap is cut into value slices of 2^kSpSliceShift values
// (32 KB of bits); a route pass sorts the query's records by slice, and one
// work item per (query, slice) ORs that slice's records into LDS and writes
// the slice once.  Queries beyond the context's record budget run dense
// instead; indices with more slices or records than the references encode
// take the atomic path (zero-fill, then one global atomic per touched word).
constexpr uint32_t kSpSliceShift = 18;
constexpr uint32_t kSpSliceWords = 1u << (kSpSliceShift - 5);
constexpr uint32_t kSpMaxSlices = 1024;
constexpr uint32_t kSpTypeShift = 29;  // ref = type << 29 | record index (< 2^29)
enum SparseFlag : uint8_t { kSpNone = 0, kSpAtomic = 1, kSpSlice = 2 };

// Arguments of the split, route and emit stages.
struct SparseArgs {
    const uint32_t *keys, *kq;  // keys, query of each key
    const uint64_t *qoff;       // nq+1 key offsets per query
    const uint64_t *koff;       // record prefix over keys (K+1)
    uint64_t K;
    const uint64_t *Kd;         // when set, K is read here on the device
    uint32_t nq, chunk;         // queries; records per route chunk (<= 65535)
    uint32_t nslices;           // value slices per bitmap
    uint8_t *flag;              // per query: kSpNone (dense) / kSpAtomic / kSpSlice
    uint32_t *cpre;             // nq+1: route-chunk prefix over queries
    uint32_t *scpre;            // nq: slice-chunk prefix (slice queries)
    uint64_t *rbase;            // nq: first ref of a slice query
    uint32_t *slist;            // slice queries in order
    uint32_t *tot;              // [0] route chunks, [1] slice queries, [2] slice chunks
    uint32_t *refs;             // routed records, sorted by slice within each chunk
    uint16_t *choff;            // per slice chunk: nslices+1 offsets into its refs
    uint64_t cap;               // refs capacity (0: every sparse query atomic)
    uint32_t *out;
    uint64_t out_stride;
};
// Device-dispatched batch, one work-group: the sparse queries' path (slice,
// atomic, or dense past the record budget; M from the features) with their
// route-chunk prefixes and refs bases, then the dense groups.
void launch_assign_groups(uint8_t *sel, const uint64_t *M, uint32_t nq, uint32_t G, uint32_t *qgroup,
                          uint32_t *grp_q, uint32_t *grp_n, uint32_t *plan, const SparseArgs &a, hipStream_t s);
// Route chunks: slice queries' records sorted into refs; atomic queries'
// records OR-ed into their (zeroed) bitmaps.  Persistent over `grid`.
hipError_t launch_sparse_route(const DevIndex &ix, const SparseArgs &a, uint32_t grid, hipStream_t s);
uint32_t sparse_route_lds_bytes(uint32_t chunk);

// Arguments of the persistent union kernel.
struct FusedArgs {
    const uint32_t *plan;        // [0] dense groups
    const uint32_t *grp_q;       // 64 query slots per group
    const uint32_t *grp_n;       // queries per group
    const void *masks;           // per-group key masks (u32 or u64 per key)
    uint64_t mask_stride;        // elements per group
    const uint32_t *anybms;      // per-group 1-bit-per-key filters
    uint64_t any_stride;         // words per group
    uint32_t *out;
    uint64_t out_stride, out_words;
    int use_col;
    // work counters, reset by the launcher: [0] the slice items, [1 + g] the
    // tiles of dense group g (ngroups_cap of them)
    uint32_t *counter;
    uint32_t ngroups_cap;
    // write-once sparse: one emit item per (slice query, slice), after the dense items
    const uint32_t *sp_tot, *sp_slist, *sp_cpre, *sp_scpre, *sp_refs;
    const uint64_t *sp_rbase;
    const uint16_t *sp_choff;
    uint32_t sp_chunk, nslices;
    // compaction chunk counts of the slice queries (ccnt NULL: none): each
    // slice item writes its query's counts of the chunks it covers to
    // ccnt[q * nchunk + chunk]
    uint32_t *ccnt;
    uint32_t nchunk;
};
uint32_t fused_lds_bytes(int mode, uint32_t G, uint32_t tile_shift, uint32_t threads, bool use_col);
// Work-groups of the persistent union launch: CUs x resident work-groups per
// CU (256 or 512 threads).  Queried before any stream capture.
hipError_t fused_grid(int mode, uint32_t threads, const DevIndex &ix, uint32_t G, bool use_col, uint32_t *grid);
// Resets the work counter and launches `grid` work-groups.
hipError_t launch_union_fused(int mode, uint32_t threads, const DevIndex &ix, const FusedArgs &a, uint32_t G,
                              uint32_t grid, hipStream_t s);
constexpr uint32_t kCompactWords = 2048;
// skip (when not NULL): queries with skip[q] == kSpSlice are left alone (the
// union's slice items counted them).
void launch_popcount(const uint32_t *bm, uint64_t stride, uint64_t words, uint32_t nq,
                     uint32_t *ccnt, hipStream_t s, const uint8_t *skip = nullptr);
// Query q's ids go to ids + id_off[q] (device offsets) or, with id_off NULL,
// to ids + q * ids_stride; off_aligned: every id_off[q] is a multiple of 4.
// ccnt (from launch_popcount) is turned into per-query exclusive prefixes in
// place (counts[q] = the totals, when counts is not NULL); ids NULL: counts only.
void launch_write_ids(const uint32_t *bm, uint64_t stride, uint64_t words, uint32_t nq,
                      uint32_t *ccnt, uint32_t *ids, uint64_t ids_stride, const uint64_t *id_off,
                      bool off_aligned, uint32_t *counts, hipStream_t s);
// Id-only batch by value slice (after launch_sparse_route has routed every
// query's records into refs/choff as slice queries): item (qi, s) ORs slice
// s of query qi in LDS, counts it and writes its ids to ids + id_off[q0 + qi]
// after the query's earlier slices; counts[q0 + qi] = the total.  st: nq *
// nslices status words and counter: zero at launch; err: set if a look-back
// ever gave up.
struct IdsSliceArgs {
    const uint32_t *cpre;   // nq+1: route-chunk prefix over the part's queries
    const uint64_t *rbase;  // nq: first ref of each query
    const uint32_t *refs;
    const uint16_t *choff;  // per route chunk: nslices+1 offsets into its refs
    uint32_t chunk, nslices, nq, q0;
    const uint64_t *id_off;
    uint32_t *ids, *counts;
    unsigned long long *st;
    uint32_t *counter, *err;
};
uint32_t ids_slices_grid();
hipError_t launch_ids_slices(const DevIndex &ix, const IdsSliceArgs &a, bool aligned, uint32_t grid, hipStream_t s);
void launch_features(const DevIndex &ix, const uint32_t *keys, const uint64_t *qoff, uint32_t nq,
                     uint64_t R, float *feats, uint64_t *M, hipStream_t s);
void launch_decide_threshold(const uint64_t *M, uint32_t nq, double thr, uint8_t *sel, hipStream_t s);

}  // namespace ab
