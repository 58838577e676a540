// gpu_build.hip -- the index build on the device: the same two layouts as
// the host build (host_index.cpp), byte for byte, from device-resident pairs.
//
//   1. pack (key, value) into u64 and radix-sort, drop duplicates;
//   2. a record starts where the key or the 256-value window changes; an
//      exclusive scan numbers the records, another (six counters wide) ranks
//      each record within its type (GroupKeys, src/AdaBool/inverted_list.h:
//      390-470: types by window size <=1, <=5, <=9, <=17, <=33, else BITSET);
//   3. one thread per record writes its header and payload; the seg8 rows
//      (key_segments_ prefixes) are the type ranks at each key's first
//      record, written without atomics;
//   4. value-tiled copy: per type a stable radix sort of records by tile, a
//      gather, then blocks of 1024 records per (tile, type), their delta
//      widths (the ComputeKeyBlockMeta rule) and byte offsets by scans, and
//      one thread per record writing its transposed delta bytes.
// Sorts and scans are rocPRIM's; everything else is kernels here.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"

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

struct Cnt6 {
    uint32_t c[6];
};
struct AddCnt6 {
    __host__ __device__ Cnt6 operator()(const Cnt6 &a, const Cnt6 &b) const
    {
        Cnt6 r;
        for (int i = 0; i < 6; ++i) r.c[i] = a.c[i] + b.c[i];
        return r;
    }
};

__host__ __device__ inline int bucket_type(uint32_t n)
{
    return n <= 1 ? AB_ARRAY1 : n <= 5 ? AB_ARRAY4 : n <= 9 ? AB_ARRAY8 : n <= 17 ? AB_ARRAY16 : n <= 33 ? AB_ARRAY32
                                                                                                        : AB_BITSET;
}

struct Packed {
    const uint64_t *kv;
    int vbits;
    uint64_t vmask;
    __device__ uint32_t key(uint64_t i) const { return (uint32_t)(vbits ? (kv[i] >> vbits) : kv[i]); }
    __device__ uint32_t val(uint64_t i) const { return (uint32_t)(kv[i] & vmask); }
};

inline uint32_t grid_for(uint64_t n, uint32_t bs = 256)
{
    uint64_t g = (n + bs - 1) / bs;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 20));
}

// Largest key and value. A bounded grid strides over the pairs, and each
// 256-thread work-group makes one pair of atomics. With a work-group per 256
// pairs, two same-address atomics per wave serialised at the L2: 95 ms for
// 4e8 pairs, against about 1 ms of reading.
constexpr uint32_t kMaxPairsGrid = 2048;
__global__ void __launch_bounds__(256) k_max_pairs(const uint32_t *k, const uint32_t *v, uint64_t n, uint32_t *out)
{
    __shared__ uint32_t s_ab[2][4];
    uint32_t a = 0, b = 0;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint32_t k0 = k[i], k1 = k[i + stride], k2 = k[i + 2 * stride], k3 = k[i + 3 * stride];
        const uint32_t v0 = v[i], v1 = v[i + stride], v2 = v[i + 2 * stride], v3 = v[i + 3 * stride];
        a = max(a, max(max(k0, k1), max(k2, k3)));
        b = max(b, max(max(v0, v1), max(v2, v3)));
    }
    for (; i < n; i += stride) {
        a = max(a, k[i]);
        b = max(b, v[i]);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        a = max(a, (uint32_t)__shfl_xor(a, d, 64));
        b = max(b, (uint32_t)__shfl_xor(b, d, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        s_ab[0][threadIdx.x >> 6] = a;
        s_ab[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = max(max(s_ab[0][0], s_ab[0][1]), max(s_ab[0][2], s_ab[0][3]));
        b = max(max(s_ab[1][0], s_ab[1][1]), max(s_ab[1][2], s_ab[1][3]));
        atomicMax(&out[0], a);
        atomicMax(&out[1], b);
    }
}

__global__ void k_pack(const uint32_t *k, const uint32_t *v, uint64_t n, int vbits, uint64_t *kv)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        kv[i] = ((uint64_t)k[i] << vbits) | v[i];
}

// 1 where a record (key, window) starts
__global__ void k_rec_flags(Packed p, uint64_t n, uint32_t *flag)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        flag[i] = (i == 0 || p.key(i) != p.key(i - 1) || (p.val(i) >> 8) != (p.val(i - 1) >> 8)) ? 1u : 0u;
}

__global__ void k_rec_starts(const uint32_t *flag, const uint32_t *rid, uint64_t n, uint32_t *start)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (flag[i]) start[rid[i]] = (uint32_t)i;
}

__global__ void k_rec_types(const uint32_t *start, uint32_t R, uint32_t n, Cnt6 *onehot)
{
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = r + 1 < R ? start[r + 1] : n;
        Cnt6 c{};
        c.c[bucket_type(e - start[r])] = 1;
        onehot[r] = c;
    }
}

struct TypeOut {
    uint32_t *hdr[6];
    uint8_t *pay[6];
    uint32_t *rkey[6];
};

// header, payload and key of every record
__global__ void k_rec_fill(Packed p, const uint32_t *start, const Cnt6 *rank, uint32_t R, uint32_t n, TypeOut o)
{
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = start[r], e = r + 1 < R ? start[r + 1] : n, cnt = e - s;
        const int t = bucket_type(cnt);
        const uint32_t idx = rank[r].c[t];
        const uint32_t key = p.key(s), v0 = p.val(s), win = v0 & 0xffffff00u;
        o.rkey[t][idx] = key;
        if (t == AB_BITSET) {
            o.hdr[t][idx] = win;
            uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (uint32_t b = s; b < e; ++b) {
                const uint32_t lo = p.val(b) & 0xffu;
                w[lo >> 5] |= 1u << (lo & 31);
            }
            uint32_t *dst = reinterpret_cast<uint32_t *>(o.pay[t] + (uint64_t)idx * 32);
            for (int i = 0; i < 8; ++i) dst[i] = w[i];
        } else {
            o.hdr[t][idx] = win | (v0 & 0xffu);
            const int pb = t == AB_ARRAY4 ? 4 : t == AB_ARRAY8 ? 8 : t == AB_ARRAY16 ? 16 : t == AB_ARRAY32 ? 32 : 0;
            uint8_t *dst = o.pay[t] + (uint64_t)idx * pb;
            for (uint32_t b = 1; b < cnt; ++b) dst[b - 1] = (uint8_t)(p.val(s + b) & 0xffu);
        }
    }
}

// seg8 rows without atomics: records are sorted by key, so row k (the counts
// of keys < k) is the state before the first record whose key is >= k: its
// six type ranks, its record index and its first pair index.  The thread of
// record r writes the rows of keys (key of record r-1, key of record r]; the
// thread of r = R writes the rows after the last key, up to row nkeys, with
// the totals.
__global__ void k_seg8_rows(Packed p, const uint32_t *start, const Cnt6 *rank, uint32_t R, uint32_t n, Cnt6 tot,
                            uint32_t nkeys, uint32_t *seg8)
{
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= R; r += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t kcur = r < R ? (int64_t)p.key(start[r]) : (int64_t)nkeys;
        const int64_t kprev = r > 0 ? (int64_t)p.key(start[r - 1]) : -1;
        if (kcur == kprev) continue;
        uint32_t row[8];
        for (int t = 0; t < 6; ++t) row[t] = r < R ? rank[r].c[t] : tot.c[t];
        row[6] = (uint32_t)r;
        row[7] = r < R ? start[r] : n;
        for (int64_t k = kprev + 1; k <= kcur; ++k) {
            uint4 *dst = reinterpret_cast<uint4 *>(seg8 + (uint64_t)k * kSegStride);
            dst[0] = make_uint4(row[0], row[1], row[2], row[3]);
            dst[1] = make_uint4(row[4], row[5], row[6], row[7]);
        }
    }
}

__global__ void k_tile_of(const uint32_t *hdr, uint32_t nr, uint32_t ts, uint32_t *tile, uint32_t *idx)
{
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nr; r += (uint64_t)gridDim.x * blockDim.x) {
        tile[r] = hdr[r] >> ts;
        idx[r] = (uint32_t)r;
    }
}

__global__ void k_gather(const uint32_t *perm, uint32_t nr, const uint32_t *hdr, const uint32_t *rkey,
                         const uint8_t *pay, int pb, uint32_t *vhdr, uint32_t *vkey, uint8_t *vpay)
{
    for (uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; d < nr; d += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = perm[d];
        vhdr[d] = hdr[r];
        vkey[d] = rkey[r];
        for (int i = 0; i < pb; ++i) vpay[d * pb + i] = pay[(uint64_t)r * pb + i];
    }
}

// tstart[j] = first record of tile j in the tile-sorted order (j = 0..ntiles)
__global__ void k_tile_starts(const uint32_t *stile, uint32_t nr, uint32_t ntiles, uint32_t *tstart)
{
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= ntiles;
         j += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nr;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (stile[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        tstart[j] = lo;
    }
}

struct VtIn {
    const uint32_t *tstart[6];
    const uint32_t *vkey[6];
};

// blocks per (tile, type), flattened f = tile * 6 + type
__global__ void k_block_counts(VtIn v, uint32_t ntiles, uint32_t *nb)
{
    for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < (uint64_t)ntiles * 6;
         f += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(f / 6), t = (uint32_t)(f % 6);
        const uint32_t c = v.tstart[t][j + 1] - v.tstart[t][j];
        nb[f] = (c + kBlockRecs - 1) / kBlockRecs;
    }
}

__device__ inline uint32_t width_of(uint32_t span) { return span < 256u ? 1 : (span < 65536u ? 2 : 4); }

// per block: its (tile, type) by search over the flattened block prefix,
// first record, record count and delta width
__global__ void k_block_meta(VtIn v, const uint32_t *bbase, uint32_t nf, uint32_t NB, uint4 *meta,
                             uint64_t *bytes)
{
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < NB; b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nf;  // last f with bbase[f] <= b
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (bbase[mid] <= b) lo = mid;
            else hi = mid;
        }
        const uint32_t f = lo, j = f / 6, t = f % 6;
        const uint32_t s0 = v.tstart[t][j], e = v.tstart[t][j + 1];
        const uint32_t s = s0 + (uint32_t)(b - bbase[f]) * kBlockRecs;
        const uint32_t nv = min(kBlockRecs, e - s);
        const uint32_t k0 = v.vkey[t][s];
        const uint32_t nxt = (s + kBlockRecs < e) ? v.vkey[t][s + kBlockRecs] : v.vkey[t][e - 1];
        const uint32_t w = width_of(nxt - k0);
        const uint32_t wl = w == 1 ? 0 : (w == 2 ? 1 : 2);
        meta[b] = make_uint4(k0, 0, t | (wl << 4) | (nv << 8), s);
        bytes[b] = (uint64_t)kBlockRecs * w;
    }
}

// one work-group per block: byte offsets into the metadata, then every
// record's delta bytes at (lane * 16 + slot) * width
__global__ __launch_bounds__(256) void k_block_keys(VtIn v, uint4 *meta, const uint64_t *boff, uint8_t *vkeys)
{
    const uint32_t b = blockIdx.x;
    uint4 m = meta[b];
    if (threadIdx.x == 0) {
        m.y = (uint32_t)(boff[b] >> 10);  // in units of 1024 bytes
        meta[b] = m;
    }
    const uint32_t t = m.z & 15u, wl = (m.z >> 4) & 3u, nv = m.z >> 8, w = 1u << wl;
    uint8_t *dst = vkeys + boff[b];
    for (uint32_t r = threadIdx.x; r < nv; r += 256) {
        const uint32_t d = v.vkey[t][m.w + r] - m.x;
        uint8_t *p = dst + ((uint64_t)(r & 63) * kSlots + (r >> 6)) * w;
        for (uint32_t q = 0; q < w; ++q) p[q] = (uint8_t)(d >> (8 * q));
    }
}

// device buffer owned until handed to the index
// (freed only after the stream has drained: every phase ends in a synchronize)
struct Tmp {
    void *p = nullptr;
    Tmp() = default;
    Tmp(const Tmp &) = delete;
    Tmp &operator=(const Tmp &) = delete;
    ~Tmp() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    hipError_t alloc(size_t bytes)
    {
        reset();
        return hipMalloc(&p, std::max<size_t>(bytes, 16));
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

int bits_for(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

}  // namespace

// Allocation of a resident array the kernels read: +64 zero bytes past the
// end (vector loads never run past it), as for uploaded host arrays.
static ab_status resident(ab_index *ix, size_t bytes, void **out)
{
    const size_t alloc = (bytes + 64 + 255) / 256 * 256;
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, alloc));
    ix->allocs.push_back(p);
    HIP_TRY(hipMemset(p, 0, alloc));
    ix->dev_bytes += alloc;
    *out = p;
    return AB_OK;
}

extern "C" ab_status ab_index_build_device(const uint32_t *d_keys, const uint32_t *d_vals, uint64_t nnz,
                                           uint32_t nkeys, uint32_t nvalues, uint32_t tile_shift, void *stream,
                                           ab_index **out)
{
    if (!out || (nnz && (!d_keys || !d_vals))) {
        set_error("ab_index_build_device: NULL argument");
        return AB_INVALID_ARG;
    }
    if (tile_shift != 0 && (tile_shift < 8 || tile_shift > 16)) {
        set_error("ab_index_build_device: tile_shift must be 0 or 8..16");
        return AB_INVALID_ARG;
    }
    if (nnz >= 0xffffffffull) {
        set_error("ab_index_build_device: at most 2^32-2 pairs");
        return AB_UNSUPPORTED;
    }
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    std::unique_ptr<ab_index> ix(new ab_index());
    HIP_TRY(hipGetDevice(&ix->device));
    // ---- key / value space (as the host build) -------------------------------
    uint64_t need_k = 0, need_v = 0;
    if (nnz) {
        Tmp mx;
        HIP_TRY(mx.alloc(8));
        HIP_TRY(hipMemsetAsync(mx.p, 0, 8, s));
        hipLaunchKernelGGL(k_max_pairs, dim3(std::min(grid_for(nnz), kMaxPairsGrid)), dim3(256), 0, s, d_keys, d_vals,
                           nnz, mx.as<uint32_t>());
        uint32_t h[2];
        HIP_TRY(hipMemcpyAsync(h, mx.p, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        need_k = (uint64_t)h[0] + 1;
        need_v = (uint64_t)h[1] + 1;
    }
    if (need_k > 0xffffffffull || need_v > 0xffffffffull) {
        set_error("ab_index_build_device: key or value 0xffffffff is not representable");
        return AB_UNSUPPORTED;
    }
    if (nkeys == 0) nkeys = (uint32_t)need_k;
    if (nvalues == 0) nvalues = (uint32_t)need_v;
    if (nkeys < need_k || nvalues < need_v) {
        set_error("ab_index_build_device: nkeys/nvalues smaller than max+1 of the pairs");
        return AB_INVALID_ARG;
    }
    const int vbits = bits_for(nvalues ? nvalues - 1 : 0), kbits = bits_for(nkeys ? nkeys - 1 : 0);
    if (vbits + kbits > 64) {
        set_error("ab_index_build_device: key and value spaces too large to pack");
        return AB_UNSUPPORTED;
    }
    const uint32_t ts = tile_shift ? tile_shift : auto_tile_shift(nvalues);
    const uint32_t ntiles = (uint32_t)(((uint64_t)nvalues + (1ull << ts) - 1) >> ts);
    const uint64_t vmask = vbits ? ((vbits == 64) ? ~0ull : ((1ull << vbits) - 1)) : 0;

    // ---- sort and deduplicate ------------------------------------------------
    Tmp kvA, kvB, cntU, tmp;
    size_t tb = 0;
    uint64_t n = 0;
    if (nnz) {
        HIP_TRY(kvA.alloc(nnz * 8));
        HIP_TRY(kvB.alloc(nnz * 8));
        hipLaunchKernelGGL(k_pack, dim3(grid_for(nnz)), dim3(256), 0, s, d_keys, d_vals, nnz, vbits,
                           kvA.as<uint64_t>());
        HIP_TRY(rocprim::radix_sort_keys(nullptr, tb, kvA.as<uint64_t>(), kvB.as<uint64_t>(), (size_t)nnz, 0,
                                         std::max(1, vbits + kbits), s));
        size_t tb2 = 0;
        HIP_TRY(cntU.alloc(8));
        HIP_TRY(rocprim::unique(nullptr, tb2, kvB.as<uint64_t>(), kvA.as<uint64_t>(), cntU.as<uint64_t>(),
                                (size_t)nnz, rocprim::equal_to<uint64_t>(), s));
        HIP_TRY(tmp.alloc(std::max(tb, tb2)));
        tb = std::max(tb, tb2);
        HIP_TRY(rocprim::radix_sort_keys(tmp.p, tb, kvA.as<uint64_t>(), kvB.as<uint64_t>(), (size_t)nnz, 0,
                                         std::max(1, vbits + kbits), s));
        HIP_TRY(rocprim::unique(tmp.p, tb, kvB.as<uint64_t>(), kvA.as<uint64_t>(), cntU.as<uint64_t>(), (size_t)nnz,
                                rocprim::equal_to<uint64_t>(), s));
        HIP_TRY(hipMemcpyAsync(&n, cntU.p, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    kvB.reset();
    const Packed P{kvA.as<uint64_t>(), vbits, vmask};

    // ---- records and their per-type ranks ----------------------------------------
    Tmp flag, rid, start, onehot, rank;
    uint32_t R = 0;
    uint32_t tot[6] = {0, 0, 0, 0, 0, 0};
    if (n) {
        HIP_TRY(flag.alloc(n * 4));
        HIP_TRY(rid.alloc(n * 4));
        hipLaunchKernelGGL(k_rec_flags, dim3(grid_for(n)), dim3(256), 0, s, P, n, flag.as<uint32_t>());
        size_t need = 0;
        HIP_TRY(rocprim::exclusive_scan(nullptr, need, flag.as<uint32_t>(), rid.as<uint32_t>(), 0u, (size_t)n,
                                        rocprim::plus<uint32_t>(), s));
        Tmp t1;
        HIP_TRY(t1.alloc(need));
        HIP_TRY(rocprim::exclusive_scan(t1.p, need, flag.as<uint32_t>(), rid.as<uint32_t>(), 0u, (size_t)n,
                                        rocprim::plus<uint32_t>(), s));
        uint32_t last[2];
        HIP_TRY(hipMemcpyAsync(&last[0], rid.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&last[1], flag.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        R = last[0] + last[1];
        HIP_TRY(start.alloc((size_t)R * 4));
        hipLaunchKernelGGL(k_rec_starts, dim3(grid_for(n)), dim3(256), 0, s, flag.as<uint32_t>(), rid.as<uint32_t>(),
                           n, start.as<uint32_t>());
        HIP_TRY(onehot.alloc((size_t)R * sizeof(Cnt6)));
        HIP_TRY(rank.alloc((size_t)R * sizeof(Cnt6)));
        hipLaunchKernelGGL(k_rec_types, dim3(grid_for(R)), dim3(256), 0, s, start.as<uint32_t>(), R, (uint32_t)n,
                           onehot.as<Cnt6>());
        const Cnt6 zero{};
        need = 0;
        HIP_TRY(rocprim::exclusive_scan(nullptr, need, onehot.as<Cnt6>(), rank.as<Cnt6>(), zero, (size_t)R, AddCnt6(),
                                        s));
        Tmp t2;
        HIP_TRY(t2.alloc(need));
        HIP_TRY(rocprim::exclusive_scan(t2.p, need, onehot.as<Cnt6>(), rank.as<Cnt6>(), zero, (size_t)R, AddCnt6(),
                                        s));
        Cnt6 lr, lo;
        HIP_TRY(hipMemcpyAsync(&lr, rank.as<Cnt6>() + R - 1, sizeof(Cnt6), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&lo, onehot.as<Cnt6>() + R - 1, sizeof(Cnt6), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int t = 0; t < 6; ++t) tot[t] = lr.c[t] + lo.c[t];
    }
    flag.reset();
    rid.reset();
    onehot.reset();

    // ---- key-major arrays and seg8 ---------------------------------------------
    DevIndex &d = ix->dev;
    TypeOut o{};
    Tmp rkey[6];
    for (int t = 0; t < 6; ++t) {
        void *h = nullptr, *p = nullptr;
        AB_TRY(resident(ix.get(), (size_t)tot[t] * 4, &h));
        AB_TRY(resident(ix.get(), (size_t)tot[t] * kPayloadBytes[t], &p));
        d.hdr[t] = (const uint32_t *)h;
        d.pay[t] = (const uint8_t *)p;
        o.hdr[t] = (uint32_t *)h;
        o.pay[t] = (uint8_t *)p;
        HIP_TRY(rkey[t].alloc((size_t)tot[t] * 4));
        o.rkey[t] = rkey[t].as<uint32_t>();
    }
    void *seg = nullptr;
    const size_t seg_words = ((size_t)nkeys + 1) * kSegStride;
    AB_TRY(resident(ix.get(), seg_words * 4, &seg));
    d.seg8 = (const uint32_t *)seg;
    if (R)
        hipLaunchKernelGGL(k_rec_fill, dim3(grid_for(R)), dim3(256), 0, s, P, start.as<uint32_t>(), rank.as<Cnt6>(), R,
                           (uint32_t)n, o);
    {
        Cnt6 totc{};
        for (int t = 0; t < 6; ++t) totc.c[t] = (uint32_t)tot[t];
        hipLaunchKernelGGL(k_seg8_rows, dim3(grid_for((uint64_t)R + 1)), dim3(256), 0, s, P, start.as<uint32_t>(),
                           rank.as<Cnt6>(), R, (uint32_t)n, totc, nkeys, (uint32_t *)seg);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
    }
    start.reset();
    rank.reset();
    kvA.reset();

    // ---- value-tiled copy ----------------------------------------------------------
    VtIn vin{};
    Tmp tstart[6], vkeyT[6];
    const int tbits = std::max(1, bits_for(ntiles ? ntiles - 1 : 0));
    for (int t = 0; t < 6; ++t) {
        const uint32_t nr = tot[t];
        const int pb = kPayloadBytes[t];
        void *vh = nullptr, *vp = nullptr;
        AB_TRY(resident(ix.get(), (size_t)nr * 4, &vh));
        AB_TRY(resident(ix.get(), (size_t)nr * pb, &vp));
        d.vhdr[t] = (const uint32_t *)vh;
        d.vpay[t] = (const uint8_t *)vp;
        HIP_TRY(vkeyT[t].alloc((size_t)nr * 4));
        HIP_TRY(tstart[t].alloc(((size_t)ntiles + 1) * 4));
        if (nr) {
            Tmp tile, idx, stile, perm;
            HIP_TRY(tile.alloc((size_t)nr * 4));
            HIP_TRY(idx.alloc((size_t)nr * 4));
            HIP_TRY(stile.alloc((size_t)nr * 4));
            HIP_TRY(perm.alloc((size_t)nr * 4));
            hipLaunchKernelGGL(k_tile_of, dim3(grid_for(nr)), dim3(256), 0, s, o.hdr[t], nr, ts, tile.as<uint32_t>(),
                               idx.as<uint32_t>());
            size_t need = 0;
            HIP_TRY(rocprim::radix_sort_pairs(nullptr, need, tile.as<uint32_t>(), stile.as<uint32_t>(),
                                              idx.as<uint32_t>(), perm.as<uint32_t>(), (size_t)nr, 0, tbits, s));
            Tmp t4;
            HIP_TRY(t4.alloc(need));
            HIP_TRY(rocprim::radix_sort_pairs(t4.p, need, tile.as<uint32_t>(), stile.as<uint32_t>(),
                                              idx.as<uint32_t>(), perm.as<uint32_t>(), (size_t)nr, 0, tbits, s));
            hipLaunchKernelGGL(k_gather, dim3(grid_for(nr)), dim3(256), 0, s, perm.as<uint32_t>(), nr, o.hdr[t],
                               o.rkey[t], o.pay[t], pb, (uint32_t *)vh, vkeyT[t].as<uint32_t>(), (uint8_t *)vp);
            hipLaunchKernelGGL(k_tile_starts, dim3(grid_for((uint64_t)ntiles + 1)), dim3(256), 0, s,
                               stile.as<uint32_t>(), nr, ntiles, tstart[t].as<uint32_t>());
            HIP_TRY(hipStreamSynchronize(s));  // the temporaries are freed at scope end
        } else {
            HIP_TRY(hipMemsetAsync(tstart[t].p, 0, ((size_t)ntiles + 1) * 4, s));
        }
        vin.tstart[t] = tstart[t].as<uint32_t>();
        vin.vkey[t] = vkeyT[t].as<uint32_t>();
    }
    for (int t = 0; t < 6; ++t) rkey[t].reset();
    // blocks
    const uint32_t nf = ntiles * 6;
    Tmp nb, bbase;
    HIP_TRY(nb.alloc(((size_t)nf + 1) * 4));
    HIP_TRY(bbase.alloc(((size_t)nf + 1) * 4));
    uint32_t NB = 0;
    if (nf) {
        hipLaunchKernelGGL(k_block_counts, dim3(grid_for(nf)), dim3(256), 0, s, vin, ntiles, nb.as<uint32_t>());
        size_t need = 0;
        HIP_TRY(rocprim::exclusive_scan(nullptr, need, nb.as<uint32_t>(), bbase.as<uint32_t>(), 0u, (size_t)nf,
                                        rocprim::plus<uint32_t>(), s));
        Tmp t5;
        HIP_TRY(t5.alloc(need));
        HIP_TRY(rocprim::exclusive_scan(t5.p, need, nb.as<uint32_t>(), bbase.as<uint32_t>(), 0u, (size_t)nf,
                                        rocprim::plus<uint32_t>(), s));
        uint32_t l[2];
        HIP_TRY(hipMemcpyAsync(&l[0], bbase.as<uint32_t>() + nf - 1, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&l[1], nb.as<uint32_t>() + nf - 1, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        NB = l[0] + l[1];
    }
    void *vb = nullptr, *tb_res = nullptr;
    AB_TRY(resident(ix.get(), (size_t)NB * 16, &vb));
    d.vblk = (const uint4 *)vb;
    // tile_blk[j] = bbase[j * 6], tile_blk[ntiles] = NB
    AB_TRY(resident(ix.get(), ((size_t)ntiles + 1) * 4, &tb_res));
    d.tile_blk = (const uint32_t *)tb_res;
    if (ntiles) {
        HIP_TRY(hipMemcpy2DAsync(tb_res, 4, bbase.p, 24, 4, ntiles, hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipMemcpyAsync((uint32_t *)tb_res + ntiles, &NB, 4, hipMemcpyHostToDevice, s));
    uint64_t kbytes = 0;
    Tmp bytes, boff;
    if (NB) {
        HIP_TRY(bytes.alloc((size_t)NB * 8));
        HIP_TRY(boff.alloc((size_t)NB * 8));
        hipLaunchKernelGGL(k_block_meta, dim3(grid_for(NB)), dim3(256), 0, s, vin, bbase.as<uint32_t>(), nf, NB,
                           (uint4 *)vb, bytes.as<uint64_t>());
        size_t need = 0;
        HIP_TRY(rocprim::exclusive_scan(nullptr, need, bytes.as<uint64_t>(), boff.as<uint64_t>(), (uint64_t)0,
                                        (size_t)NB, rocprim::plus<uint64_t>(), s));
        Tmp t6;
        HIP_TRY(t6.alloc(need));
        HIP_TRY(rocprim::exclusive_scan(t6.p, need, bytes.as<uint64_t>(), boff.as<uint64_t>(), (uint64_t)0,
                                        (size_t)NB, rocprim::plus<uint64_t>(), s));
        uint64_t l[2];
        HIP_TRY(hipMemcpyAsync(&l[0], boff.as<uint64_t>() + NB - 1, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(&l[1], bytes.as<uint64_t>() + NB - 1, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        kbytes = l[0] + l[1];
    }
    if ((kbytes >> 10) >= 0xffffffffull) {
        set_error("ab_index_build_device: key stream exceeds 4 TiB");
        return AB_UNSUPPORTED;
    }
    void *vk = nullptr;
    AB_TRY(resident(ix.get(), (size_t)kbytes + 64, &vk));
    d.vkeys = (const uint8_t *)vk;
    if (NB)
        hipLaunchKernelGGL(k_block_keys, dim3(NB), dim3(256), 0, s, vin, (uint4 *)vb, boff.as<uint64_t>(),
                           (uint8_t *)vk);
    HIP_TRY(hipGetLastError());
    // ---- host-side metadata and the seg8 copy host features use ----------------------
    ix->seg8.resize(seg_words);
    HIP_TRY(hipMemcpyAsync(ix->seg8.data(), seg, seg_words * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    d.nkeys = nkeys;
    d.nvalues = nvalues;
    d.tile_shift = ts;
    d.ntiles = ntiles;
    ix->nkeys = nkeys;
    ix->nvalues = nvalues;
    ix->nnz = n;
    ix->R = 0;
    ix->list_bytes = 0;
    for (int t = 0; t < 6; ++t) {
        ix->nrec[t] = tot[t];
        ix->R += tot[t];
        ix->list_bytes += (uint64_t)tot[t] * (4 + kPayloadBytes[t]);
    }
    ix->tile_shift = ts;
    ix->ntiles = ntiles;
    ix->nblocks = NB;
    ix->kbytes = kbytes;
    ix->stride = stride_bytes(nvalues);
    ix->host_only = false;
    *out = ix.release();
    return AB_OK;
}
