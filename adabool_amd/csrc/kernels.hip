// kernels.hip -- CDNA4 (gfx950) kernels of the AdaBool OR-filtering path.
//
// Record emission follows ProcessArrayFunctor (src/AdaBool/inverted_list.cu.h:
// 61-94) and the bitset emit (src/AdaBool/key_transformer.cu.h:100-111):
// header value always; payload bytes [0, kMinNum) always, later bytes until the
// first zero; a bitset ORs its 8 words into window words (header >> 5) + j.
// Instead of one atomic per value, the values of a record (sorted, all in one
// 256-value window) are folded into per-word bit masks and each touched word
// is OR-ed once.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "adabool.h"
#include "device.h"
#include "kernels.h"

namespace ab {

// ---------------------------------------------------------------------------
// fills
// ---------------------------------------------------------------------------
__global__ void k_fill_zero(uint4 *p, uint64_t n16)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void k_fill_zero_tail(uint8_t *p, uint64_t n)
{
    uint64_t i = threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_fill_zero(void *p, uint64_t bytes, hipStream_t s)
{
    if (!bytes) return;
    uint64_t n16 = bytes / 16;
    if (n16) {
        uint64_t grid = (n16 + 255) / 256;
        if (grid > 4096) grid = 4096;
        hipLaunchKernelGGL(k_fill_zero, dim3((uint32_t)grid), dim3(256), 0, s, (uint4 *)p, n16);
    }
    if (bytes % 16)
        hipLaunchKernelGGL(k_fill_zero_tail, dim3(1), dim3(64), 0, s, (uint8_t *)p + n16 * 16, bytes % 16);
}

constexpr uint32_t kZeroChunk = 8192;  // words per work-group
constexpr uint32_t kNoGroup = 0xffffffffu;
__global__ __launch_bounds__(256) void k_zero_qlist(uint32_t *bm, uint64_t stride, uint64_t words,
                                                    const uint32_t *qlist)
{
    const uint32_t q = qlist[blockIdx.y];
    const uint64_t w0 = (uint64_t)blockIdx.x * kZeroChunk;
    uint32_t *dst = bm + q * stride;
    const uint64_t w1 = w0 + kZeroChunk < words ? w0 + kZeroChunk : words;
    for (uint64_t i = w0 + threadIdx.x; i < w1; i += 256) dst[i] = 0;
}

void launch_zero_qlist(uint32_t *bm, uint64_t stride, uint64_t words, const uint32_t *qlist, uint32_t n,
                       hipStream_t s)
{
    if (!n || !words) return;
    uint32_t gx = (uint32_t)((words + kZeroChunk - 1) / kZeroChunk);
    for (uint32_t i0 = 0; i0 < n; i0 += 65535) {  // the list goes on the grid's y dimension
        const uint32_t m = n - i0 < 65535 ? n - i0 : 65535;
        hipLaunchKernelGGL(k_zero_qlist, dim3(gx, m), dim3(256), 0, s, bm, stride, words, qlist + i0);
    }
}

// ---------------------------------------------------------------------------
// exclusive scan of u64 counts: out[0..n] (out[n] = total)
// ---------------------------------------------------------------------------
constexpr uint32_t kScanNT = 256, kScanIPT = 16, kScanTile = kScanNT * kScanIPT;

uint64_t scan_parts_needed(uint64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// Block-wide exclusive scan of one value per thread; *total gets the sum.
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long x, unsigned long long *sh,
                                                              unsigned long long *total)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    unsigned long long incl = wave_incl_scan64(x);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    unsigned long long before = 0, all = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        if (w < wave) before += sh[w];
        all += sh[w];
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

__global__ __launch_bounds__(kScanNT) void k_scan_tile_sums(const uint64_t *in, uint64_t n, const uint64_t *n_dev,
                                                             uint64_t *parts)
{
    __shared__ unsigned long long sh[kScanNT / 64];
    if (n_dev) n = *n_dev;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    unsigned long long s = 0;
    for (uint32_t i = 0; i < kScanIPT; ++i) {
        uint64_t j = base + (uint64_t)i * kScanNT + threadIdx.x;
        if (j < n) s += in[j];
    }
    unsigned long long tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) parts[blockIdx.x] = tot;
}

// single work-group: exclusive scan of parts[0..np) in place; parts[np] = total
__global__ __launch_bounds__(kScanNT) void k_scan_parts(uint64_t *parts, uint64_t np, const uint64_t *n_dev)
{
    __shared__ unsigned long long sh[kScanNT / 64];
    if (n_dev) np = (*n_dev + kScanTile - 1) / kScanTile;
    unsigned long long carry = 0;
    for (uint64_t base = 0; base < np; base += kScanNT) {
        uint64_t j = base + threadIdx.x;
        unsigned long long x = j < np ? parts[j] : 0, tot;
        unsigned long long e = block_excl_scan(x, sh, &tot);
        if (j < np) parts[j] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) parts[np] = carry;
}

__global__ __launch_bounds__(kScanNT) void k_scan_apply(const uint64_t *in, uint64_t n, const uint64_t *n_dev,
                                                         const uint64_t *parts, uint64_t np, uint64_t *out)
{
    __shared__ unsigned long long sh[kScanNT / 64];
    if (n_dev) {
        n = *n_dev;
        np = (n + kScanTile - 1) / kScanTile;
        if (np == 0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
            return;
        }
        if (blockIdx.x >= np) return;  // block-uniform
    }
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanIPT;
    unsigned long long v[kScanIPT], s = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanIPT; ++i) {
        uint64_t j = base + i;
        v[i] = j < n ? in[j] : 0;
        s += v[i];
    }
    unsigned long long tot;
    unsigned long long e = block_excl_scan(s, sh, &tot) + parts[blockIdx.x];
#pragma unroll
    for (uint32_t i = 0; i < kScanIPT; ++i) {
        uint64_t j = base + i;
        if (j < n) out[j] = e;
        e += v[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = parts[np];
}

void launch_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *parts, hipStream_t s,
                     const uint64_t *n_dev)
{
    uint64_t np = (n + kScanTile - 1) / kScanTile;
    if (np == 0) {
        // out[0] = 0
        launch_fill_zero(out, 8, s);
        return;
    }
    hipLaunchKernelGGL(k_scan_tile_sums, dim3((uint32_t)np), dim3(kScanNT), 0, s, in, n, n_dev, parts);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanNT), 0, s, parts, np, n_dev);
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)np), dim3(kScanNT), 0, s, in, n, n_dev, (const uint64_t *)parts,
                       np, out);
}

// ---------------------------------------------------------------------------
// sparse path: plan + load-balanced list union (key-major layout)
// ---------------------------------------------------------------------------
// Per key j of query q: cnt[j] = records of key j (all types) if q takes the
// sparse path and the key is in range (Prepare(keys) extents,
// src/AdaBool/inverted_list.h:491-505), else 0.
__global__ __launch_bounds__(256) void k_sparse_plan(DevIndex ix, const uint32_t *__restrict__ keys,
                                                     const uint64_t *__restrict__ qoff,
                                                     const uint8_t *__restrict__ sel, uint32_t *kq, uint64_t *cnt)
{
    const uint32_t q = blockIdx.x;
    const uint64_t a = qoff[q], b = qoff[q + 1];
    const bool on = sel == nullptr || sel[q] == AB_PATH_SPARSE;
    for (uint64_t j = a + threadIdx.x; j < b; j += 256) {
        const uint32_t k = keys[j];
        uint64_t c = 0;
        if (on && k < ix.nkeys) c = ix.seg8[(uint64_t)(k + 1) * 8 + 6] - ix.seg8[(uint64_t)k * 8 + 6];
        cnt[j] = c;
        kq[j] = q;
    }
}

void launch_sparse_plan(const DevIndex &ix, const uint32_t *keys, const uint64_t *qoff, const uint8_t *sel,
                        uint32_t nq, uint32_t *kq, uint64_t *cnt, hipStream_t s)
{
    if (!nq) return;
    hipLaunchKernelGGL(k_sparse_plan, dim3(nq), dim3(256), 0, s, ix, keys, qoff, sel, kq, cnt);
}

template <int T>
__device__ __forceinline__ void emit_global(const DevIndex &ix, uint32_t idx, uint32_t *__restrict__ bm)
{
    const uint32_t h = ix.hdr[T][idx];
    uint32_t w[8];
    load_payload<T>(ix.pay[T], idx, w);
    emit_record<T>(h, w, [&](uint32_t word, uint32_t bits) { atomicOr(bm + word, bits); });
}

// Work item = one record of one (query, key).  Items are numbered through the
// prefix koff over keys; each work-group takes CH consecutive items (merge-path
// split of the total over all lists, instead of the reference's CPU chunking
// into <=256-record units, src/AdaBool/load_balance_transformer.h:50-69).
constexpr uint32_t kSparseNT = 256, kSparseWin = 1024;

// Records [c0, c1) of the sparse work: every thread of the work-group takes
// items c0 + tid, c0 + tid + NT, ...; skoff is an LDS window of kSparseWin+1
// key offsets, s_lo/s_hi LDS scratch of the search.  Calls
// f(std::integral_constant<int, T>, idx, it, q) for record idx of type T (the
// key-major arrays), item number it, of query q.
template <int NT, class F>
__device__ __forceinline__ void sparse_records(const DevIndex &ix, const uint32_t *__restrict__ keys,
                                               const uint32_t *__restrict__ kq, const uint64_t *__restrict__ koff,
                                               uint64_t K, uint64_t c0, uint64_t c1, unsigned long long *skoff,
                                               unsigned long long *s_lo, unsigned long long *s_hi, F &&f)
{
    uint64_t jw = wg_find<NT>(koff, K, c0, s_lo, s_hi);
    for (;;) {
        const uint64_t jn = (jw + kSparseWin < K) ? jw + kSparseWin : K;
        const uint32_t nwin = (uint32_t)(jn - jw);
        for (uint32_t i = threadIdx.x; i <= nwin; i += NT) skoff[i] = koff[jw + i];
        __syncthreads();
        const uint64_t wlo = skoff[0] > c0 ? skoff[0] : c0;
        const uint64_t whi = skoff[nwin] < c1 ? skoff[nwin] : c1;
        for (uint64_t it = wlo + threadIdx.x; it < whi; it += NT) {
            uint32_t lo = 0, hi = nwin;  // skoff[lo] <= it < skoff[hi]
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (skoff[mid] <= it) lo = mid; else hi = mid;
            }
            const uint64_t j = jw + lo;
            uint32_t r = (uint32_t)(it - skoff[lo]);
            const uint32_t k = keys[j];
            const uint32_t q = kq[j];
            const uint4 *rows = reinterpret_cast<const uint4 *>(ix.seg8 + (uint64_t)k * 8);
            const uint4 a0 = rows[0], a1 = rows[1], b0 = rows[2], b1 = rows[3];
            const uint32_t l0 = b0.x - a0.x, l1 = b0.y - a0.y, l2 = b0.z - a0.z, l3 = b0.w - a0.w,
                           l4 = b1.x - a1.x;
            if (r < l0) { f(std::integral_constant<int, AB_ARRAY1>{}, a0.x + r, it, q); continue; }
            r -= l0;
            if (r < l1) { f(std::integral_constant<int, AB_ARRAY4>{}, a0.y + r, it, q); continue; }
            r -= l1;
            if (r < l2) { f(std::integral_constant<int, AB_ARRAY8>{}, a0.z + r, it, q); continue; }
            r -= l2;
            if (r < l3) { f(std::integral_constant<int, AB_ARRAY16>{}, a0.w + r, it, q); continue; }
            r -= l3;
            if (r < l4) { f(std::integral_constant<int, AB_ARRAY32>{}, a1.x + r, it, q); continue; }
            r -= l4;
            f(std::integral_constant<int, AB_BITSET>{}, a1.y + r, it, q);
        }
        const bool done = (skoff[nwin] >= c1) || jn >= K;
        __syncthreads();
        if (done) break;
        jw = jn;
    }
}

// The records [c0, c1), each OR-ed into its query's bitmap with one global
// atomic per touched word (the bitmaps are zeroed beforehand).
template <int NT>
__device__ __forceinline__ void sparse_chunk(const DevIndex &ix, const uint32_t *__restrict__ keys,
                                             const uint32_t *__restrict__ kq, const uint64_t *__restrict__ koff,
                                             uint64_t K, uint64_t c0, uint64_t c1, uint32_t *__restrict__ out,
                                             uint64_t out_stride, unsigned long long *skoff,
                                             unsigned long long *s_lo, unsigned long long *s_hi)
{
    sparse_records<NT>(ix, keys, kq, koff, K, c0, c1, skoff, s_lo, s_hi,
                       [&](auto t, uint32_t idx, uint64_t, uint32_t q) {
                           emit_global<decltype(t)::value>(ix, idx, out + (uint64_t)q * out_stride);
                       });
}

__global__ __launch_bounds__(kSparseNT) void k_sparse_union(DevIndex ix, const uint32_t *__restrict__ keys,
                                                            const uint32_t *__restrict__ kq,
                                                            const uint64_t *__restrict__ koff, uint64_t K,
                                                            uint32_t *__restrict__ out, uint64_t out_stride,
                                                            uint32_t CH)
{
    __shared__ unsigned long long skoff[kSparseWin + 1];
    __shared__ unsigned long long s_lo, s_hi;
    if (K == 0) return;
    const uint64_t W = koff[K];
    const uint64_t nch = (W + CH - 1) / CH;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint64_t c0 = c * CH, c1 = (c0 + CH < W) ? c0 + CH : W;
        sparse_chunk<kSparseNT>(ix, keys, kq, koff, K, c0, c1, out, out_stride, skoff, &s_lo, &s_hi);
    }
}

void launch_sparse_union(const DevIndex &ix, const uint32_t *keys, const uint32_t *kq, const uint64_t *koff,
                         uint64_t K, uint32_t *out, uint64_t out_stride, uint32_t chunk, uint32_t grid,
                         hipStream_t s)
{
    if (!K || !grid) return;
    hipLaunchKernelGGL(k_sparse_union, dim3(grid), dim3(kSparseNT), 0, s, ix, keys, kq, koff, K, out, out_stride,
                       chunk);
}

// ---------------------------------------------------------------------------
// dense path: query masks
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void k_mask_keys(const uint32_t *__restrict__ keys, const uint64_t *__restrict__ qoff,
                                                   DenseGroup g, void *mask, uint32_t *anybm, uint32_t nkeys,
                                                   int clear)
{
    const uint32_t b = blockIdx.y;
    const uint32_t q = g.q[b];
    const uint64_t a = qoff[q], e = qoff[q + 1];
    for (uint64_t j = a + (uint64_t)blockIdx.x * 256 + threadIdx.x; j < e; j += (uint64_t)gridDim.x * 256) {
        const uint32_t k = keys[j];
        if (k >= nkeys) continue;
        // the 1-bit-per-key bitmap: the query itself in bitmap mode, the
        // union of the group's queries (a filter in front of mask) otherwise
        uint32_t *w = anybm + (k >> 5);
        if (clear) *w = 0; else atomicOr(w, 1u << (k & 31));
        if constexpr (MODE == kModeMask32) {
            uint32_t *m = reinterpret_cast<uint32_t *>(mask) + k;
            if (clear) *m = 0; else atomicOr(m, 1u << b);
        } else if constexpr (MODE == kModeMask64) {
            unsigned long long *m = reinterpret_cast<unsigned long long *>(mask) + k;
            if (clear) *m = 0; else atomicOr(m, 1ull << b);
        }
    }
}

void launch_mask_keys(int mode, const uint32_t *keys, const uint64_t *qoff, const DenseGroup &g, void *mask,
                      uint32_t *anybm, uint32_t nkeys, int clear, hipStream_t s)
{
    dim3 grid(8, g.B);
    if (mode == kModeBitmap)
        hipLaunchKernelGGL(k_mask_keys<kModeBitmap>, grid, dim3(256), 0, s, keys, qoff, g, mask, anybm, nkeys, clear);
    else if (mode == kModeMask32)
        hipLaunchKernelGGL(k_mask_keys<kModeMask32>, grid, dim3(256), 0, s, keys, qoff, g, mask, anybm, nkeys, clear);
    else
        hipLaunchKernelGGL(k_mask_keys<kModeMask64>, grid, dim3(256), 0, s, keys, qoff, g, mask, anybm, nkeys, clear);
}

// mask[k] bit b = bit k of bitmap (b0 + b), and anybm bit k = (mask[k] != 0),
// for every k < nkeys (every entry written; waves cover 64 aligned keys).
__global__ __launch_bounds__(256) void k_mask_from_bitmaps(const uint32_t *__restrict__ bms, uint64_t len_u32,
                                                           uint32_t b0, uint32_t B, unsigned long long *mask,
                                                           uint32_t *anybm, uint32_t nkeys)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u));
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t kb = wave0; kb < nkeys; kb += stride) {
        const uint64_t k = kb + lane;
        unsigned long long m = 0;
        if (k < nkeys) {
            const uint64_t w = k >> 5;
            if (w < len_u32)
                for (uint32_t b = 0; b < B; ++b)
                    m |= (unsigned long long)((bms[(uint64_t)(b0 + b) * len_u32 + w] >> (k & 31)) & 1u) << b;
            mask[k] = m;
        }
        const unsigned long long any = __ballot(m != 0);
        if (lane == 0) anybm[kb >> 5] = (uint32_t)any;
        if (lane == 32 && kb + 32 < nkeys) anybm[(kb >> 5) + 1] = (uint32_t)(any >> 32);
    }
}

void launch_mask_from_bitmaps(const uint32_t *bitmaps, uint64_t len_u32, uint32_t b0, uint32_t B, void *mask,
                              uint32_t *anybm, uint32_t nkeys, hipStream_t s)
{
    uint64_t grid = ((uint64_t)nkeys + 255) / 256;
    if (grid > 8192) grid = 8192;
    if (!grid) return;
    hipLaunchKernelGGL(k_mask_from_bitmaps, dim3((uint32_t)grid), dim3(256), 0, s, bitmaps, len_u32, b0, B,
                       (unsigned long long *)mask, anybm, nkeys);
}

// ---------------------------------------------------------------------------
// dense path: value-tiled bitmap probe, one work-group per value tile
// ---------------------------------------------------------------------------
// the query set of a key (or of a value, in the column accumulators): one bit
// per query of the group
struct M128 {
    unsigned long long lo, hi;
};
template <int MODE> struct MaskOf { using T = uint32_t; };
template <> struct MaskOf<kModeMask64> { using T = unsigned long long; };
template <> struct MaskOf<kModeMask128> { using T = M128; };
struct M256 {
    unsigned long long w[4];
};
template <> struct MaskOf<kModeMask256> { using T = M256; };

__device__ __forceinline__ bool any_bits(uint32_t m) { return m != 0; }
__device__ __forceinline__ bool any_bits(unsigned long long m) { return m != 0; }
__device__ __forceinline__ bool any_bits(const M128 &m) { return (m.lo | m.hi) != 0; }
__device__ __forceinline__ bool any_bits(const M256 &m) { return (m.w[0] | m.w[1] | m.w[2] | m.w[3]) != 0; }
__device__ __forceinline__ uint32_t popc_mask(uint32_t m) { return (uint32_t)__popc(m); }
__device__ __forceinline__ uint32_t popc_mask(unsigned long long m) { return (uint32_t)__popcll(m); }
__device__ __forceinline__ uint32_t popc_mask(const M128 &m) { return (uint32_t)(__popcll(m.lo) + __popcll(m.hi)); }
template <class F> __device__ __forceinline__ void for_bits(unsigned long long m, uint32_t base, F &&f)
{
    while (m) {
        f(base + (uint32_t)__builtin_ctzll(m));
        m &= m - 1;
    }
}
template <class F> __device__ __forceinline__ void for_bits(uint32_t m, F &&f) { for_bits((unsigned long long)m, 0, f); }
template <class F> __device__ __forceinline__ void for_bits(unsigned long long m, F &&f) { for_bits(m, 0, f); }
template <class F> __device__ __forceinline__ void for_bits(const M128 &m, F &&f)
{
    for_bits(m.lo, 0, f);
    for_bits(m.hi, 64, f);
}
template <class F> __device__ __forceinline__ void for_bits(const M256 &m, F &&f)
{
    for_bits(m.w[0], 0, f);
    for_bits(m.w[1], 64, f);
    for_bits(m.w[2], 128, f);
    for_bits(m.w[3], 192, f);
}
// 64 x 64 bit transpose across a wave: lane l holds row l (bit c = M[l][c]);
// afterwards lane b holds column b (bit l = M[l][b]).  Six butterfly stages,
// each swapping the off-diagonal s x s blocks between lanes l and l ^ s; the
// partner's word comes by DPP or ds_swizzle where the pattern allows, and by
// ds_bpermute only for lane ^ 32.
__device__ __forceinline__ unsigned long long transpose64(unsigned long long x, uint32_t lane)
{
    constexpr unsigned long long kLow[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                            0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t s = 32u >> k;
        unsigned long long y;
        if (s == 32) {
            y = __shfl_xor(x, 32, 64);
        } else if (s == 16 || s == 4) {
            // xor inside 32 lanes: ds_swizzle in bit-mask mode (and 0x1f, xor s)
            const int lo = (int)(uint32_t)x, hi = (int)(uint32_t)(x >> 32);
            int ylo, yhi;
            if (s == 16) {
                ylo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
                yhi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
            } else {
                ylo = __builtin_amdgcn_ds_swizzle(lo, 0x101F);
                yhi = __builtin_amdgcn_ds_swizzle(hi, 0x101F);
            }
            y = ((unsigned long long)(uint32_t)yhi << 32) | (uint32_t)ylo;
        } else if (s == 8) {
            // lane ^ 8 is a rotation by 8 inside a row of 16: DPP row_ror:8
            const int lo = (int)(uint32_t)x, hi = (int)(uint32_t)(x >> 32);
            const int ylo = __builtin_amdgcn_mov_dpp(lo, 0x128, 0xF, 0xF, false);
            const int yhi = __builtin_amdgcn_mov_dpp(hi, 0x128, 0xF, 0xF, false);
            y = ((unsigned long long)(uint32_t)yhi << 32) | (uint32_t)ylo;
        } else {
            // lane ^ 1 and lane ^ 2 stay inside a quad: DPP quad permutes
            constexpr int kQuad1 = 0xB1, kQuad2 = 0x4E;  // [1,0,3,2], [2,3,0,1]
            const int lo = (int)(uint32_t)x, hi = (int)(uint32_t)(x >> 32);
            int ylo, yhi;
            if (s == 1) {
                ylo = __builtin_amdgcn_mov_dpp(lo, kQuad1, 0xF, 0xF, false);
                yhi = __builtin_amdgcn_mov_dpp(hi, kQuad1, 0xF, 0xF, false);
            } else {
                ylo = __builtin_amdgcn_mov_dpp(lo, kQuad2, 0xF, 0xF, false);
                yhi = __builtin_amdgcn_mov_dpp(hi, kQuad2, 0xF, 0xF, false);
            }
            y = ((unsigned long long)(uint32_t)yhi << 32) | (uint32_t)ylo;
        }
        x = (lane & s) ? ((x & ~kLow[k]) | ((y & ~kLow[k]) >> s)) : ((x & kLow[k]) | ((y & kLow[k]) << s));
    }
    return x;
}
// Column masks in LDS (one query set per value of the tile, tv values): a
// mask of up to 64 bits is one word per value; wider ones are kept as planes
// of u64 words (word c of value v at plane c, index c * tv + v), so that the
// lanes of a wave ORing into different values hit 32 bank pairs instead of
// the 8 that a 16- or 32-byte stride per value leaves them.
__device__ __forceinline__ void col_or(uint32_t *col, uint32_t v, uint32_t, uint32_t m) { atomicOr(&col[v], m); }
__device__ __forceinline__ void col_or(unsigned long long *col, uint32_t v, uint32_t, unsigned long long m)
{
    atomicOr(&col[v], m);
}
// both halves (or all four words) unconditionally: lanes of a wave hold
// different masks, and a test per word would split the wave into partly
// masked atomics anyway (ORing a zero word is harmless; 5% less union time
// than testing each half)
__device__ __forceinline__ void col_or(M128 *col, uint32_t v, uint32_t tv, const M128 &m)
{
    unsigned long long *cp = reinterpret_cast<unsigned long long *>(col);
    atomicOr(&cp[v], m.lo);
    atomicOr(&cp[tv + v], m.hi);
}
__device__ __forceinline__ void col_or(M256 *col, uint32_t v, uint32_t tv, const M256 &m)
{
    unsigned long long *cp = reinterpret_cast<unsigned long long *>(col);
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicOr(&cp[i * tv + v], m.w[i]);
}
// 64-query chunk c of the column mask of value v
__device__ __forceinline__ unsigned long long col_word(const uint32_t *col, uint32_t v, uint32_t, uint32_t)
{
    return col[v];
}
__device__ __forceinline__ unsigned long long col_word(const unsigned long long *col, uint32_t v, uint32_t, uint32_t)
{
    return col[v];
}
template <class M>
__device__ __forceinline__ unsigned long long col_word(const M *col, uint32_t v, uint32_t tv, uint32_t c)
{
    return reinterpret_cast<const unsigned long long *>(col)[c * tv + v];
}
// 64-query chunk c of a mask
__device__ __forceinline__ unsigned long long mask_word(uint32_t m, uint32_t) { return m; }
__device__ __forceinline__ unsigned long long mask_word(unsigned long long m, uint32_t) { return m; }
__device__ __forceinline__ unsigned long long mask_word(const M128 &m, uint32_t c) { return c ? m.hi : m.lo; }
__device__ __forceinline__ unsigned long long mask_word(const M256 &m, uint32_t c)
{
    return c == 0 ? m.w[0] : (c == 1 ? m.w[1] : (c == 2 ? m.w[2] : m.w[3]));
}
// bit b of a mask in memory (global atomics)
__device__ __forceinline__ void set_mask_bit(uint32_t *p, uint32_t b) { atomicOr(p, 1u << b); }
__device__ __forceinline__ void set_mask_bit(unsigned long long *p, uint32_t b) { atomicOr(p, 1ull << b); }
__device__ __forceinline__ void set_mask_bit(M128 *p, uint32_t b)
{
    atomicOr(b < 64 ? &p->lo : &p->hi, 1ull << (b & 63));
}
__device__ __forceinline__ void set_mask_bit(M256 *p, uint32_t b) { atomicOr(&p->w[b >> 6], 1ull << (b & 63)); }

// Query bits of key k: the 1-bit-per-key bitmap answers alone for one query
// and filters the per-key mask loads (8 bytes per key) for a group.
template <int MODE>
__device__ __forceinline__ typename MaskOf<MODE>::T probe(const void *__restrict__ m,
                                                          const uint32_t *__restrict__ anybm, uint32_t k)
{
    const uint32_t hit = (anybm[k >> 5] >> (k & 31)) & 1u;
    if constexpr (MODE == kModeBitmap)
        return hit;
    else if constexpr (MODE == kModeMask32)
        return hit ? reinterpret_cast<const uint32_t *>(m)[k] : 0u;
    else
        return hit ? reinterpret_cast<const unsigned long long *>(m)[k] : 0ull;
}

// Rank words of device-dispatched groups (one uint2 per 32 keys: the
// any-query bits, the rank of the first key): the u32 holding key k's bit,
// and key k's rank among the group's keys.
__device__ __forceinline__ uint32_t rank_word_index(uint32_t k) { return (k >> 5) * 2; }
__device__ __forceinline__ uint32_t key_rank(const uint2 &w, uint32_t k)
{
    return w.y + (uint32_t)__popc(w.x & ((1u << (k & 31)) - 1));
}

// Picks w[j] for a run-time j < N without dynamic register indexing.
template <int N>
__device__ __forceinline__ uint32_t pick(const uint32_t (&w)[8], uint32_t j)
{
    uint32_t r = w[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = (j == (uint32_t)i) ? w[i] : r;
    return r;
}

// ORs one matched record into the LDS tiles of every query whose bit is set in
// m.  Two accumulators: row tiles (one bit row per query; a record costs one
// OR per touched word per query) and, for query groups, a column mask per
// value (the set of queries containing that value; a record costs one OR per
// value); the write-out transposes the columns into the rows with wave
// ballots.  When a group has column masks, every record uses them: choosing
// per record the cheaper accumulator split the lanes of a wave between both
// code paths, and the uniform choice took the C2 union from 3.00 to 2.81 ms.
// Without columns, the record's values are sorted and share one 256-value
// window, so runs of values in one word fold into a single row OR.  The
// payload is walked 4 bytes per step to keep one copy of the code.
template <int MODE, int T>
__device__ __forceinline__ void dense_emit(uint32_t h, const uint32_t (&w)[8], typename MaskOf<MODE>::T m,
                                           uint32_t *tiles, uint32_t tw, uint32_t tile_word0,
                                           typename MaskOf<MODE>::T *col, uint32_t vmask)
{
    auto flush = [&](uint32_t gw, uint32_t bits) {
        const uint32_t lw = gw - tile_word0;
        if constexpr (MODE == kModeBitmap) {
            atomicOr(&tiles[lw], bits);
        } else {
            for_bits(m, [&](uint32_t b) { atomicOr(&tiles[b * tw + lw], bits); });
        }
    };
    auto put = [&](uint32_t v) { col_or(col, v & vmask, vmask + 1, m); };
    const bool use_col = MODE != kModeBitmap && col != nullptr;
    if constexpr (T == AB_BITSET) {
        const uint32_t base = h >> 5;
        // words are taken from the front of a rotating copy: no run-time
        // register indexing (which the compiler would put in scratch memory)
        uint32_t ww[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ww[i] = w[i];
#pragma unroll 1
        for (uint32_t j = 0; j < 8; ++j) {
            uint32_t x = ww[0];
#pragma unroll
            for (int i = 0; i < 7; ++i) ww[i] = ww[i + 1];
            if (!x) continue;
            if constexpr (MODE != kModeBitmap) {
                if (use_col) {
                    while (x) {
                        put((h & 0xffffff00u) + j * 32 + (uint32_t)__builtin_ctz(x));
                        x &= x - 1;
                    }
                    continue;
                }
            }
            flush(base + j, x);
        }
    } else if constexpr (T == AB_ARRAY1) {
        if (use_col) {
            put(h);
            return;
        }
        flush(h >> 5, 1u << (h & 31));
    } else {
        constexpr int P = TypeTraits<T>::P, kMin = TypeTraits<T>::kMin;
        const uint32_t hi = h & 0xffffff00u;
        uint32_t cw = h >> 5, bits = 1u << (h & 31);
        if (use_col) put(h);
        bool stop = false;
        uint32_t ww[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ww[i] = w[i];
#pragma unroll 1
        for (uint32_t j = 0; j < (uint32_t)(P / 4) && !stop; ++j) {
            const uint32_t word = ww[0];
#pragma unroll
            for (int i = 0; i < P / 4 - 1; ++i) ww[i] = ww[i + 1];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t b = (word >> (8 * k)) & 0xffu;
                if (!stop && (j * 4 + k >= (uint32_t)kMin) && b == 0) stop = true;
                if (!stop) {
                    const uint32_t v = hi | b;
                    if (use_col) {
                        put(v);
                    } else {
                        if ((v >> 5) != cw) {
                            flush(cw, bits);
                            cw = v >> 5;
                            bits = 0;
                        }
                        bits |= 1u << (v & 31);
                    }
                }
            }
        }
        if (!use_col) flush(cw, bits);
    }
}

// One 1024-record block of type T whose absolute keys are staged in kb
// (record s*64 + lane at kb[s*64 + lane]).  Slots go RS at a time (eight for
// ARRAY1, whose round fits in few registers; four for the others, which
// measured 1% faster than eight for ARRAY4/8 once every record used the
// column masks): the RS any-bit probes first; then, for every slot whose key some
// query of the group holds, the query-mask load and the header/payload loads
// together (the any-bit is exactly "mask != 0"); then the RS emits -- unrolled
// for ARRAY1, whose emit is a few instructions, else taken from the front of a
// rotating register set (no run-time register indexing, so no scratch).
template <int MODE, int T, bool RANK>
__device__ __forceinline__ void dense_block(const DevIndex &ix, const uint32_t *kb, uint32_t nv, uint32_t rec_base,
                                            const void *__restrict__ mask, const uint32_t *__restrict__ anybm,
                                            uint32_t lane, uint32_t *tiles, uint32_t tw, uint32_t tile_word0,
                                            typename MaskOf<MODE>::T *col, uint32_t vmask)
{
    using M = typename MaskOf<MODE>::T;
    constexpr int RS = (T == AB_ARRAY1) ? 8 : 4;
    constexpr int PW = TypeTraits<T>::P / 4 > 0 ? TypeTraits<T>::P / 4 : 1;  // payload words kept per slot
    const uint32_t *__restrict__ hdr = ix.vhdr[T];
    const uint8_t *__restrict__ pay = ix.vpay[T];
#pragma unroll 1
    for (uint32_t s0 = 0; s0 < kSlots; s0 += RS) {
        uint32_t key[RS], hit[RS];
        uint32_t any_hit = 0;
#pragma unroll
        for (int j = 0; j < RS; ++j) {
            const uint32_t r = (s0 + j) * 64 + lane;
            key[j] = r < nv ? kb[r] : 0u;
            if constexpr (RANK) {
                // one 8-byte load: the key's any-query bit and its rank in the
                // group's mask table (key[j] becomes the mask index)
                const uint2 rw = reinterpret_cast<const uint2 *>(anybm)[key[j] >> 5];
                hit[j] = r < nv ? (rw.x >> (key[j] & 31)) & 1u : 0u;
                key[j] = key_rank(rw, key[j]);
            } else {
                hit[j] = r < nv ? (anybm[key[j] >> 5] >> (key[j] & 31)) & 1u : 0u;
            }
            any_hit |= hit[j];
        }
        if (!any_hit) continue;
        M m[RS];
        uint32_t h[RS], w[RS][8];
#pragma unroll
        for (int j = 0; j < RS; ++j) {
            m[j] = M{};
            h[j] = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) w[j][i] = 0;
            if (hit[j]) {
                if constexpr (MODE == kModeBitmap)
                    m[j] = 1u;
                else
                    m[j] = reinterpret_cast<const M *>(mask)[key[j]];
                const uint32_t idx = rec_base + (s0 + j) * 64 + lane;
                h[j] = hdr[idx];
                load_payload<T>(pay, idx, w[j]);
            }
        }
        if constexpr (T == AB_ARRAY1) {
#pragma unroll
            for (int j = 0; j < RS; ++j)
                if (any_bits(m[j])) dense_emit<MODE, T>(h[j], w[j], m[j], tiles, tw, tile_word0, col, vmask);
        } else {
#pragma unroll 1
            for (uint32_t j = 0; j < (uint32_t)RS; ++j) {
                if (any_bits(m[0])) dense_emit<MODE, T>(h[0], w[0], m[0], tiles, tw, tile_word0, col, vmask);
#pragma unroll
                for (int q = 0; q < RS - 1; ++q) {
                    m[q] = m[q + 1];
                    h[q] = h[q + 1];
#pragma unroll
                    for (int i = 0; i < PW; ++i) w[q][i] = w[q + 1][i];
                }
                m[RS - 1] = M{};
            }
        }
    }
}

// One value tile for a group of B queries (output slots qids[0..B)).
// lds: per-wave key staging (1024 u32 per wave), B query row tiles, and for
// query groups with use_col one column mask per value of the tile.
template <int MODE, int NT, bool RANK = false>
__device__ __forceinline__ void dense_tile(const DevIndex &ix, uint32_t tile, const void *__restrict__ mask,
                                           const uint32_t *__restrict__ anybm, uint32_t B,
                                           const uint32_t *qids, uint32_t *__restrict__ out, uint64_t out_stride,
                                           uint64_t out_words, int use_col, uint32_t *lds)
{
    using M = typename MaskOf<MODE>::T;
    constexpr uint32_t NW = NT / 64;
    uint32_t *tiles = lds + NW * kBlockRecs;
    const uint32_t tv = 1u << ix.tile_shift, vmask = tv - 1;
    const uint32_t tw = tv >> 5;
    const uint32_t tile_word0 = tile * tw;
    const uint32_t ntw = B * tw;
    // with column masks the rows are never accumulated: the columns are
    // transposed straight into the output, so no row tiles are kept in LDS
    M *col = (MODE != kModeBitmap && use_col) ? reinterpret_cast<M *>(tiles) : nullptr;
    if (col)
        for (uint32_t i = threadIdx.x; i < tv; i += NT) col[i] = M{};
    else
        for (uint32_t i = threadIdx.x; i < ntw; i += NT) tiles[i] = 0;
    __syncthreads();

    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t *kb = lds + wave * kBlockRecs;
    const uint32_t b0 = ix.tile_blk[tile], b1 = ix.tile_blk[tile + 1];
    for (uint32_t b = b0 + wave; b < b1; b += NW) {
        const uint4 mt = ix.vblk[b];
        const uint32_t type = mt.z & 15u, wl = (mt.z >> 4) & 3u, nv = mt.z >> 8;
        const uint8_t *kp = ix.vkeys + ((uint64_t)mt.y << 10);
        // decode the block's keys (record s*64 + lane in slot s) into the wave's staging area
        if (wl == 0) {
            const uint4 v = *reinterpret_cast<const uint4 *>(kp + lane * 16u);
            const uint32_t ww[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int s = 0; s < 16; ++s) kb[s * 64 + lane] = mt.x + ((ww[s >> 2] >> (8 * (s & 3))) & 0xffu);
        } else if (wl == 1) {
            const uint4 *p = reinterpret_cast<const uint4 *>(kp + lane * 32u);
            const uint4 v0 = p[0], v1 = p[1];
            const uint32_t ww[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int s = 0; s < 16; ++s) kb[s * 64 + lane] = mt.x + ((ww[s >> 1] >> (16 * (s & 1))) & 0xffffu);
        } else {
            const uint4 *p = reinterpret_cast<const uint4 *>(kp + lane * 64u);
            const uint4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
            const uint32_t ww[16] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w,
                                     v2.x, v2.y, v2.z, v2.w, v3.x, v3.y, v3.z, v3.w};
#pragma unroll
            for (int s = 0; s < 16; ++s) kb[s * 64 + lane] = mt.x + ww[s];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        switch (type) {
        case AB_ARRAY1: dense_block<MODE, AB_ARRAY1, RANK>(ix, kb, nv, mt.w, mask, anybm, lane, tiles, tw, tile_word0, col, vmask); break;
        case AB_ARRAY4: dense_block<MODE, AB_ARRAY4, RANK>(ix, kb, nv, mt.w, mask, anybm, lane, tiles, tw, tile_word0, col, vmask); break;
        case AB_ARRAY8: dense_block<MODE, AB_ARRAY8, RANK>(ix, kb, nv, mt.w, mask, anybm, lane, tiles, tw, tile_word0, col, vmask); break;
        case AB_ARRAY16: dense_block<MODE, AB_ARRAY16, RANK>(ix, kb, nv, mt.w, mask, anybm, lane, tiles, tw, tile_word0, col, vmask); break;
        case AB_ARRAY32: dense_block<MODE, AB_ARRAY32, RANK>(ix, kb, nv, mt.w, mask, anybm, lane, tiles, tw, tile_word0, col, vmask); break;
        default: dense_block<MODE, AB_BITSET, RANK>(ix, kb, nv, mt.w, mask, anybm, lane, tiles, tw, tile_word0, col, vmask); break;
        }
        // the staging area is rewritten by the next block of this wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    const uint64_t nwq = out_words > tile_word0 ? ((out_words - tile_word0) < tw ? out_words - tile_word0 : tw) : 0;
    if (col) {
        // columns -> output rows: lane l holds the query set of value 64*gi + l;
        // a 64 x 64 bit transpose across the wave gives lane b query b's 64
        // bits of those values (64 queries at a time), two words of its row
        // that the lane stores at once.  Every word of the tile is written,
        // for every query of the group (zeros included).
        const bool pair = ((out_stride & 1) == 0) && ((reinterpret_cast<uintptr_t>(out) & 7) == 0);
        for (uint32_t gi = wave; gi < (tv >> 6); gi += NW) {
            const uint32_t w = 2 * gi;  // first of the two words, within the tile's row
            for (uint32_t c = 0; c * 64 < B; ++c) {
                const unsigned long long qv = col_word(col, gi * 64 + lane, tv, c);
                const uint32_t nb = B - c * 64 < 64 ? B - c * 64 : 64;
                unsigned long long mine = 0;
                if (__ballot(qv != 0) != 0) mine = transpose64(qv, lane);
                if (lane < nb && w < nwq) {
                    uint32_t *dst = out + (uint64_t)qids[c * 64 + lane] * out_stride + tile_word0 + w;
                    if (w + 1 < nwq) {
                        if (pair) {
                            *reinterpret_cast<uint2 *>(dst) = make_uint2((uint32_t)mine, (uint32_t)(mine >> 32));
                        } else {
                            dst[0] = (uint32_t)mine;
                            dst[1] = (uint32_t)(mine >> 32);
                        }
                    } else {
                        dst[0] = (uint32_t)mine;
                    }
                }
            }
        }
        return;
    }
    // write-once: every word of the tile, for every query of the group
    if ((out_stride & 3) == 0 && (nwq & 3) == 0) {
        const uint32_t n4 = (uint32_t)(nwq >> 2), tw4 = tw >> 2;
        const uint4 *src = reinterpret_cast<const uint4 *>(tiles);
        for (uint32_t i = threadIdx.x; i < B * n4; i += NT) {
            const uint32_t b = i / n4, k = i - b * n4;
            reinterpret_cast<uint4 *>(out + (uint64_t)qids[b] * out_stride + tile_word0)[k] = src[b * tw4 + k];
        }
    } else {
        for (uint32_t b = 0; b < B; ++b) {
            uint32_t *dst = out + (uint64_t)qids[b] * out_stride + tile_word0;
            for (uint32_t i = threadIdx.x; i < nwq; i += NT) dst[i] = tiles[b * tw + i];
        }
    }
}

template <int MODE, int NT>
__global__ __launch_bounds__(NT) void k_dense_tiles(DevIndex ix, const void *__restrict__ mask,
                                                    const uint32_t *__restrict__ anybm, DenseGroup g,
                                                    uint32_t *__restrict__ out, uint64_t out_stride,
                                                    uint64_t out_words, int use_col)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    dense_tile<MODE, NT>(ix, blockIdx.x, mask, anybm, g.B, g.q, out, out_stride, out_words, use_col, lds);
}

uint32_t dense_lds_bytes(int mode, uint32_t B, uint32_t tile_shift, uint32_t threads, bool use_col)
{
    // key staging, then either the column masks (transposed straight to the
    // output) or the row tiles
    const uint32_t staging = (threads / 64) * kBlockRecs * 4u;
    if (mode != kModeBitmap && use_col)
        return staging + (1u << tile_shift) *
                             (mode == kModeMask32 ? 4u : (mode == kModeMask64 ? 8u : (mode == kModeMask128 ? 16u : 32u)));
    return staging + B * (1u << (tile_shift - 5)) * 4u;
}

template <int MODE>
static hipError_t launch_dense_mode(uint32_t threads, const DevIndex &ix, const void *mask, const uint32_t *anybm,
                                    const DenseGroup &g, uint32_t *out, uint64_t out_stride, uint64_t out_words,
                                    bool use_col, hipStream_t s)
{
    const uint32_t lds = dense_lds_bytes(MODE, g.B, ix.tile_shift, threads, use_col);
    const int uc = use_col ? 1 : 0;
    dim3 grid(ix.ntiles);
    switch (threads) {
    case 64: hipLaunchKernelGGL((k_dense_tiles<MODE, 64>), grid, dim3(64), lds, s, ix, mask, anybm, g, out, out_stride, out_words, uc); break;
    case 128: hipLaunchKernelGGL((k_dense_tiles<MODE, 128>), grid, dim3(128), lds, s, ix, mask, anybm, g, out, out_stride, out_words, uc); break;
    case 512: hipLaunchKernelGGL((k_dense_tiles<MODE, 512>), grid, dim3(512), lds, s, ix, mask, anybm, g, out, out_stride, out_words, uc); break;
    default: hipLaunchKernelGGL((k_dense_tiles<MODE, 256>), grid, dim3(256), lds, s, ix, mask, anybm, g, out, out_stride, out_words, uc); break;
    }
    return hipGetLastError();
}

hipError_t launch_dense_tiles(int mode, uint32_t threads, const DevIndex &ix, const void *mask, const uint32_t *anybm,
                              const DenseGroup &g, uint32_t *out, uint64_t out_stride, uint64_t out_words,
                              bool use_col, hipStream_t s)
{
    if (!ix.ntiles || !g.B) return hipSuccess;
    if (mode == kModeBitmap)
        return launch_dense_mode<kModeBitmap>(threads, ix, mask, anybm, g, out, out_stride, out_words, false, s);
    if (mode == kModeMask32)
        return launch_dense_mode<kModeMask32>(threads, ix, mask, anybm, g, out, out_stride, out_words, use_col, s);
    return launch_dense_mode<kModeMask64>(threads, ix, mask, anybm, g, out, out_stride, out_words, use_col, s);
}

// ---------------------------------------------------------------------------
// device-dispatched batches: group assignment, per-group query masks, and one
// persistent union launch over dense (group, tile) items and sparse chunks
// ---------------------------------------------------------------------------
// One work-group of 1024 threads, queries in order.
// Sparse queries first: with a record budget (a.cap > 0), a sparse query takes
// the write-once slice path while the running total of sparse records M (from
// the features) fits the budget -- its refs start at that running total --
// and runs dense instead past it (sel[q] rewritten); without a budget it takes
// the atomic path.  Slice and atomic queries get ceil(M / chunk) route chunks.
// Then dense queries in query order get rank r -> group r / G, slot r % G;
// qgroup[q] = group * kGroupSlots + slot (kNoGroup otherwise); plan = {groups,
// dense}; a.tot = {route chunks, slice queries, slice chunks}.
__global__ __launch_bounds__(1024) void k_assign_groups(uint8_t *__restrict__ sel, const uint64_t *__restrict__ M,
                                                        uint32_t nq, uint32_t G, uint32_t *qgroup, uint32_t *grp_q,
                                                        uint32_t *grp_n, uint32_t *plan, SparseArgs a)
{
    __shared__ unsigned long long sh[16];
    unsigned long long run_d = 0, run_m = 0, run_ch = 0, run_sq = 0, run_sch = 0, tot;
    for (uint32_t base = 0; base < nq; base += 1024) {
        const uint32_t q = base + threadIdx.x;
        const bool in = q < nq;
        const uint8_t s0 = in ? sel[q] : (uint8_t)0;
        const bool sparse = in && s0 == AB_PATH_SPARSE;
        const unsigned long long m = (sparse && M) ? M[q] : 0;
        const unsigned long long m_before = run_m + block_excl_scan(m, sh, &tot);
        run_m += tot;
        const bool slice = sparse && a.cap && m_before + m <= a.cap;
        const bool atomic = sparse && !a.cap;
        const bool d = (in && s0 == AB_PATH_DENSE) || (sparse && !slice && !atomic);
        if (in && d && s0 != AB_PATH_DENSE) sel[q] = AB_PATH_DENSE;
        const unsigned long long nch = (slice || atomic) ? (m + a.chunk - 1) / a.chunk : 0;
        const unsigned long long ch_before = run_ch + block_excl_scan(nch, sh, &tot);
        run_ch += tot;
        const unsigned long long rank_s = run_sq + block_excl_scan(slice ? 1ull : 0ull, sh, &tot);
        run_sq += tot;
        const unsigned long long sch_before = run_sch + block_excl_scan(slice ? nch : 0ull, sh, &tot);
        run_sch += tot;
        const unsigned long long rank = run_d + block_excl_scan(d ? 1ull : 0ull, sh, &tot);
        run_d += tot;
        if (in) {
            if (d) {
                const uint32_t g = (uint32_t)(rank / G), b = (uint32_t)(rank - (unsigned long long)g * G);
                qgroup[q] = g * kGroupSlots + b;
                grp_q[g * kGroupSlots + b] = q;
            } else {
                qgroup[q] = kNoGroup;
            }
            a.flag[q] = slice ? kSpSlice : (atomic ? kSpAtomic : kSpNone);
            a.cpre[q] = (uint32_t)ch_before;
            if (slice) {
                a.scpre[q] = (uint32_t)sch_before;
                a.rbase[q] = m_before;
                a.slist[rank_s] = q;
            }
        }
    }
    const uint32_t dense = (uint32_t)run_d;
    const uint32_t ng = (dense + G - 1) / G;
    for (uint32_t g = threadIdx.x; g < ng; g += 1024) grp_n[g] = (dense - g * G) < G ? dense - g * G : G;
    if (threadIdx.x == 0) {
        plan[0] = ng;
        plan[1] = dense;
        a.cpre[nq] = (uint32_t)run_ch;
        a.tot[0] = (uint32_t)run_ch;
        a.tot[1] = (uint32_t)run_sq;
        a.tot[2] = (uint32_t)run_sch;
    }
}

void launch_assign_groups(uint8_t *sel, const uint64_t *M, uint32_t nq, uint32_t G, uint32_t *qgroup,
                          uint32_t *grp_q, uint32_t *grp_n, uint32_t *plan, const SparseArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_assign_groups, dim3(1), dim3(1024), 0, s, sel, M, nq, G, qgroup, grp_q, grp_n, plan, a);
}

__global__ __launch_bounds__(256) void k_zero_sparse(uint32_t *bm, uint64_t stride, uint64_t words,
                                                     const uint8_t *__restrict__ sel, uint8_t match)
{
    const uint32_t q = blockIdx.y;
    if (sel[q] != match) return;
    const uint64_t w0 = (uint64_t)blockIdx.x * kZeroChunk;
    const uint64_t w1 = w0 + kZeroChunk < words ? w0 + kZeroChunk : words;
    uint32_t *dst = bm + (uint64_t)q * stride;
    for (uint64_t i = w0 + threadIdx.x; i < w1; i += 256) dst[i] = 0;
}

// queries go on the grid's y dimension, in slices of at most 65535
constexpr uint32_t kMaxGridY = 65535;

void launch_zero_sparse(uint32_t *bm, uint64_t stride, uint64_t words, const uint8_t *sel, uint32_t nq,
                        uint8_t match, hipStream_t s)
{
    if (!nq || !words) return;
    const uint32_t gx = (uint32_t)((words + kZeroChunk - 1) / kZeroChunk);
    for (uint32_t q0 = 0; q0 < nq; q0 += kMaxGridY) {
        const uint32_t m = nq - q0 < kMaxGridY ? nq - q0 : kMaxGridY;
        hipLaunchKernelGGL(k_zero_sparse, dim3(gx, m), dim3(256), 0, s, bm + (uint64_t)q0 * stride, stride, words,
                           sel + q0, match);
    }
}

// Group query masks of device-dispatched batches, compacted by rank.  Per
// group, the keys its queries hold are numbered in key order (their rank);
// the masks are stored at masks[rank], so that a group's table holds only
// the keys it uses and stays small enough for the L2 cache (a table indexed
// by key is 32 bytes x nkeys per 256-query group).  The rank words hold, per
// 64 keys, the any-query bits (two u32) and the rank of the first of them,
// so that one 16-byte load gives both the probe and the rank.

// op kMaskAny: set the keys' any-query bits; kMaskSet: OR each query's bit
// into its keys' masks (after k_mask_rank); kMaskClear: zero those masks.
// One thread per key of the batch (kq: the query of each key); the key count
// comes from Kd on the device when it is set (graph replays).
template <int MODE>
__global__ __launch_bounds__(256) void k_mask_groups(const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ kq, const uint64_t *Kd,
                                                     uint64_t K, const uint32_t *__restrict__ qgroup, void *masks,
                                                     uint64_t mask_stride, uint32_t *anybms, uint64_t any_stride,
                                                     uint32_t nkeys, int op)
{
    using M = typename MaskOf<MODE>::T;
    const uint64_t n = Kd ? *Kd : K;
    for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (uint64_t)gridDim.x * 256) {
        const uint32_t k = keys[j];
        if (k >= nkeys) continue;
        const uint32_t qg = qgroup[kq[j]];
        if (qg == kNoGroup) continue;
        const uint32_t g = qg / kGroupSlots, b = qg % kGroupSlots;
        uint32_t *any = anybms + (uint64_t)g * any_stride;
        if (op == kMaskAny) {
            atomicOr(&any[rank_word_index(k)], 1u << (k & 31));
        } else {
            M *mask = reinterpret_cast<M *>(masks) + (uint64_t)g * mask_stride;
            const uint32_t r = key_rank(reinterpret_cast<const uint2 *>(any)[k >> 5], k);
            if (op == kMaskSet)
                set_mask_bit(&mask[r], b);
            else
                mask[r] = M{};
        }
    }
}

void launch_mask_groups(int mode, const uint32_t *keys, const uint32_t *kq, const uint64_t *Kd, uint64_t K,
                        const uint32_t *qgroup, void *masks, uint64_t mask_stride, uint32_t *anybms,
                        uint64_t any_stride, uint32_t nkeys, int op, hipStream_t s)
{
    if (!K) return;
    const dim3 grid((uint32_t)std::min<uint64_t>((K + 255) / 256, 8192));
    if (mode == kModeMask32)
        hipLaunchKernelGGL(k_mask_groups<kModeMask32>, grid, dim3(256), 0, s, keys, kq, Kd, K, qgroup, masks,
                           mask_stride, anybms, any_stride, nkeys, op);
    else if (mode == kModeMask64)
        hipLaunchKernelGGL(k_mask_groups<kModeMask64>, grid, dim3(256), 0, s, keys, kq, Kd, K, qgroup, masks,
                           mask_stride, anybms, any_stride, nkeys, op);
    else if (mode == kModeMask128)
        hipLaunchKernelGGL(k_mask_groups<kModeMask128>, grid, dim3(256), 0, s, keys, kq, Kd, K, qgroup, masks,
                           mask_stride, anybms, any_stride, nkeys, op);
    else
        hipLaunchKernelGGL(k_mask_groups<kModeMask256>, grid, dim3(256), 0, s, keys, kq, Kd, K, qgroup, masks,
                           mask_stride, anybms, any_stride, nkeys, op);
}

// Ranks of group g's first keys per 32, in two passes of one wave per 64 rank
// words: the exclusive prefix inside the wave's run (and the run's total),
// then the totals of the earlier runs added.  Groups past plan[0] hold no
// bits and are skipped.
__global__ __launch_bounds__(64) void k_mask_rank_local(uint32_t *anybms, uint64_t any_stride, uint32_t nw,
                                                        const uint32_t *__restrict__ plan)
{
    const uint32_t g = blockIdx.y;
    if (g >= plan[0]) return;
    uint32_t *base = anybms + (uint64_t)g * any_stride;
    uint2 *rw = reinterpret_cast<uint2 *>(base);
    uint32_t *runtot = base + ((uint64_t)nw + 1) * 2;
    const uint32_t lane = lane_id(), w = blockIdx.x * 64 + lane;
    const uint32_t n = w < nw ? (uint32_t)__popc(rw[w].x) : 0u;
    const uint32_t incl = wave_incl_scan_dpp(n);
    if (w < nw) rw[w].y = incl - n;
    if (lane == 63) runtot[blockIdx.x] = incl;
}

__global__ __launch_bounds__(64) void k_mask_rank_add(uint32_t *anybms, uint64_t any_stride, uint32_t nw,
                                                      const uint32_t *__restrict__ plan)
{
    const uint32_t g = blockIdx.y;
    if (g >= plan[0] || blockIdx.x == 0) return;
    uint32_t *base = anybms + (uint64_t)g * any_stride;
    uint2 *rw = reinterpret_cast<uint2 *>(base);
    const uint32_t *runtot = base + ((uint64_t)nw + 1) * 2;
    const uint32_t lane = lane_id(), w = blockIdx.x * 64 + lane;
    uint32_t pre = 0;
    for (uint32_t r = lane; r < blockIdx.x; r += 64) pre += runtot[r];
    pre = wave_sum(pre);
    if (w < nw) rw[w].y += pre;
}

void launch_mask_rank(uint32_t *anybms, uint64_t any_stride, uint32_t nkeys, uint32_t ngroups, const uint32_t *plan,
                      hipStream_t s)
{
    if (!ngroups) return;
    const uint32_t nw = (nkeys + 31) / 32;
    const dim3 grid((nw + 63) / 64, ngroups);
    hipLaunchKernelGGL(k_mask_rank_local, grid, dim3(64), 0, s, anybms, any_stride, nw, plan);
    hipLaunchKernelGGL(k_mask_rank_add, grid, dim3(64), 0, s, anybms, any_stride, nw, plan);
}

// ---------------------------------------------------------------------------
// write-once sparse: route and the per-(query, slice) emit (the split into
// slice, atomic and overflow-to-dense queries is made by k_assign_groups)
// ---------------------------------------------------------------------------
uint32_t sparse_route_lds_bytes(uint32_t chunk)
{
    // key-offset window | slice histogram | staged refs | staged slices
    return (kSparseWin + 1) * 8 + (kSpMaxSlices + 1) * 4 + chunk * 4 + ((chunk * 2 + 15) & ~15u);
}

// Route chunk r: query q = the last with cpre[q] <= r, records
// [qb + c * chunk, ...) of its key-prefix range.  Atomic queries: OR-ed into
// the bitmap.  Slice queries: each record's ref staged in LDS with its slice,
// counted per slice, the chunk's slice offsets written to choff, and the refs
// written slice by slice to refs[rbase[q] + c * chunk ...].
template <int NT>
__global__ __launch_bounds__(NT) void k_sparse_route(DevIndex ix, SparseArgs a)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    __shared__ unsigned long long s_lo, s_hi, sh[NT / 64];
    __shared__ uint32_t s_q;
    unsigned long long *skoff = reinterpret_cast<unsigned long long *>(lds);
    uint32_t *hist = lds + 2 * (kSparseWin + 1);
    uint32_t *st_ref = hist + (kSpMaxSlices + 1);
    uint16_t *st_sl = reinterpret_cast<uint16_t *>(st_ref + a.chunk);
    const uint64_t K = a.Kd ? *a.Kd : a.K;
    const uint32_t nr = a.tot[0];
    const uint32_t S = a.nslices;
    for (uint32_t r = blockIdx.x; r < nr; r += gridDim.x) {
        if (threadIdx.x == 0) {
            uint32_t lo = 0, hi = a.nq;  // cpre[lo] <= r < cpre[hi]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.cpre[mid] <= r) lo = mid; else hi = mid;
            }
            s_q = lo;
        }
        __syncthreads();
        const uint32_t q = s_q;
        const uint32_t c = r - a.cpre[q];
        const uint64_t qb = a.koff[a.qoff[q]], qe = a.koff[a.qoff[q + 1]];
        const uint64_t c0 = qb + (uint64_t)c * a.chunk;
        const uint64_t c1 = c0 + a.chunk < qe ? c0 + a.chunk : qe;
        if (a.flag[q] != kSpSlice) {
            sparse_chunk<NT>(ix, a.keys, a.kq, a.koff, K, c0, c1, a.out, a.out_stride, skoff, &s_lo, &s_hi);
            continue;  // sparse_chunk ends with a barrier
        }
        for (uint32_t i = threadIdx.x; i <= S; i += NT) hist[i] = 0;
        __syncthreads();
        sparse_records<NT>(ix, a.keys, a.kq, a.koff, K, c0, c1, skoff, &s_lo, &s_hi,
                           [&](auto t, uint32_t idx, uint64_t it, uint32_t) {
                               constexpr int T = decltype(t)::value;
                               const uint32_t sl = ix.hdr[T][idx] >> kSpSliceShift;
                               const uint32_t i = (uint32_t)(it - c0);
                               st_ref[i] = ((uint32_t)T << kSpTypeShift) | idx;
                               st_sl[i] = (uint16_t)sl;
                               atomicAdd(&hist[sl], 1u);
                           });
        // (sparse_records ends with a barrier)  per-slice offsets, in place
        const uint32_t per = (S + NT - 1) / NT, b0 = threadIdx.x * per;
        unsigned long long mine = 0;
        for (uint32_t b = b0; b < b0 + per && b < S; ++b) mine += hist[b];
        unsigned long long total;
        unsigned long long run = block_excl_scan(mine, sh, &total);
        uint16_t *co = a.choff + (uint64_t)(a.scpre[q] + c) * (S + 1);
        for (uint32_t b = b0; b < b0 + per && b < S; ++b) {
            const uint32_t v = hist[b];
            hist[b] = (uint32_t)run;
            co[b] = (uint16_t)run;
            run += v;
        }
        const uint32_t n = (uint32_t)(c1 - c0);
        if (threadIdx.x == 0) co[S] = (uint16_t)n;
        __syncthreads();
        uint32_t *dst = a.refs + a.rbase[q] + (uint64_t)c * a.chunk;
        for (uint32_t i = threadIdx.x; i < n; i += NT) {
            const uint32_t pos = atomicAdd(&hist[st_sl[i]], 1u);
            dst[pos] = st_ref[i];
        }
        __syncthreads();
    }
}

hipError_t launch_sparse_route(const DevIndex &ix, const SparseArgs &a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_sparse_route<kSparseNT>, dim3(grid), dim3(kSparseNT), sparse_route_lds_bytes(a.chunk), s,
                       ix, a);
    return hipGetLastError();
}

// ORs record `ref` (type << 29 | index, key-major arrays) into an LDS slice
// whose first word is w0.
__device__ __forceinline__ void emit_ref_lds(const DevIndex &ix, uint32_t ref, uint32_t *slice, uint32_t w0)
{
    const uint32_t idx = ref & ((1u << kSpTypeShift) - 1);
    // a routed record lies in its slice; the bound keeps LDS accesses inside it regardless
    auto put = [&](uint32_t word, uint32_t bits) {
        const uint32_t o = word - w0;
        if (o < kSpSliceWords) atomicOr(&slice[o], bits);
    };
    uint32_t w[8];
    switch (ref >> kSpTypeShift) {
    case AB_ARRAY1: {
        const uint32_t h = ix.hdr[AB_ARRAY1][idx];
        put(h >> 5, 1u << (h & 31));
        break;
    }
    case AB_ARRAY4: load_payload<AB_ARRAY4>(ix.pay[AB_ARRAY4], idx, w); emit_record<AB_ARRAY4>(ix.hdr[AB_ARRAY4][idx], w, put); break;
    case AB_ARRAY8: load_payload<AB_ARRAY8>(ix.pay[AB_ARRAY8], idx, w); emit_record<AB_ARRAY8>(ix.hdr[AB_ARRAY8][idx], w, put); break;
    case AB_ARRAY16: load_payload<AB_ARRAY16>(ix.pay[AB_ARRAY16], idx, w); emit_record<AB_ARRAY16>(ix.hdr[AB_ARRAY16][idx], w, put); break;
    case AB_ARRAY32: load_payload<AB_ARRAY32>(ix.pay[AB_ARRAY32], idx, w); emit_record<AB_ARRAY32>(ix.hdr[AB_ARRAY32][idx], w, put); break;
    default: load_payload<AB_BITSET>(ix.pay[AB_BITSET], idx, w); emit_record<AB_BITSET>(ix.hdr[AB_BITSET][idx], w, put); break;
    }
}

// ORs slice s of one slice query's routed records (nch route chunks of
// `chunk` references each, co: their per-chunk slice offsets) into the LDS
// slice whose first word is w0.  Up to NT chunks at a time: every thread reads
// one chunk's range of this slice, a scan places the ranges, and the threads
// take the references of all of them together (one chunk after the other cost
// a dependent load pair per chunk).  Ends with a barrier.
template <int NT>
__device__ __forceinline__ void or_slice_refs(const DevIndex &ix, const uint16_t *co, const uint32_t *refs,
                                              uint32_t nch, uint32_t S, uint32_t chunk, uint32_t s, uint32_t w0,
                                              uint32_t *slice, uint32_t *s_lo, uint32_t *s_pre,
                                              unsigned long long *s_sh)
{
    for (uint32_t j0 = 0; j0 < nch; j0 += NT) {
        const uint32_t j = j0 + threadIdx.x;
        uint32_t lo = 0, n = 0;
        if (j < nch) {
            const uint16_t *c = co + (uint64_t)j * (S + 1);
            lo = c[s];
            n = c[s + 1] - lo;
        }
        unsigned long long tot;
        const uint32_t ex = (uint32_t)block_excl_scan(n, s_sh, &tot);
        s_lo[threadIdx.x] = lo;
        s_pre[threadIdx.x] = ex;
        __syncthreads();
        const uint32_t m = nch - j0 < NT ? nch - j0 : NT;
        for (uint32_t i = threadIdx.x; i < (uint32_t)tot; i += NT) {
            uint32_t a0 = 0, a1 = m;  // the last chunk with s_pre <= i (it holds i)
            while (a1 - a0 > 1) {
                const uint32_t mid = (a0 + a1) >> 1;
                if (s_pre[mid] <= i) a0 = mid; else a1 = mid;
            }
            const uint32_t t = s_lo[a0] + (i - s_pre[a0]);
            emit_ref_lds(ix, refs[(uint64_t)(j0 + a0) * chunk + t], slice, w0);
        }
        __syncthreads();
    }
}

// Slice s of slice query q: its records from every route chunk of q, OR-ed
// into LDS, then every word of the slice written once.
template <int NT>
__device__ __forceinline__ void sparse_emit_slice(const DevIndex &ix, const FusedArgs &a, uint32_t q, uint32_t s,
                                                  uint32_t *slice)
{
    const uint64_t w0 = (uint64_t)s * kSpSliceWords;
    for (uint32_t i = threadIdx.x; i < kSpSliceWords; i += NT) slice[i] = 0;
    __syncthreads();
    const uint32_t S = a.nslices;
    const uint32_t nch = a.sp_cpre[q + 1] - a.sp_cpre[q];
    const uint16_t *co = a.sp_choff + (uint64_t)a.sp_scpre[q] * (S + 1);
    const uint32_t *refs = a.sp_refs + a.sp_rbase[q];
    __shared__ uint32_t s_lo[NT], s_pre[NT];
    __shared__ unsigned long long s_sh[NT / 64];
    or_slice_refs<NT>(ix, co, refs, nch, S, a.sp_chunk, s, (uint32_t)w0, slice, s_lo, s_pre, s_sh);
    if (w0 >= a.out_words) return;  // block-uniform
    const uint64_t nw = a.out_words - w0 < kSpSliceWords ? a.out_words - w0 : kSpSliceWords;
    if (a.ccnt) {
        // the slice's compaction chunk counts (words past the bitmap are zero)
        constexpr uint32_t CPS = kSpSliceWords / kCompactWords;
        __shared__ uint32_t s_part[CPS][NT / 64];
        for (uint32_t j = 0; j < CPS; ++j) {
            uint32_t n = 0;
            for (uint32_t i = threadIdx.x; i < kCompactWords; i += NT) n += __popc(slice[j * kCompactWords + i]);
            n = wave_sum(n);
            if (lane_id() == 0) s_part[j][threadIdx.x >> 6] = n;
        }
        __syncthreads();
        if (threadIdx.x < CPS) {
            const uint32_t c = s * CPS + threadIdx.x;
            uint32_t n = 0;
            for (uint32_t w = 0; w < NT / 64; ++w) n += s_part[threadIdx.x][w];
            if (c < a.nchunk) a.ccnt[(uint64_t)q * a.nchunk + c] = n;
        }
    }
    uint32_t *dst = a.out + (uint64_t)q * a.out_stride + w0;
    if ((a.out_stride & 3) == 0 && (nw & 3) == 0) {
        const uint4 *src = reinterpret_cast<const uint4 *>(slice);
        for (uint32_t i = threadIdx.x; i < nw / 4; i += NT) reinterpret_cast<uint4 *>(dst)[i] = src[i];
    } else {
        for (uint32_t i = threadIdx.x; i < nw; i += NT) dst[i] = slice[i];
    }
}

// Persistent: every work-group claims dense (group, tile) items -- the tiles
// of the group of its XCD first (work-groups are dispatched to the 8 XCDs in
// turn, so XCD = blockIdx.x % 8; each XCD's L2 then serves one group's rank
// words and masks), then the other groups' -- and then the (slice query,
// slice) emit items of the write-once sparse work, until none are left.
// Dense items come first: they are the long ones.  Group affinity took the
// union of a C3-mix batch of 4096 from 3.55 to 3.41 ms (C2, four groups:
// unchanged).
// Two 512-thread work-groups per CU need at most 128 VGPRs (4 waves per
// SIMD); past that only one fits and the union slows by half.
template <int MODE, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 512 ? 4 : 1)))
void k_union_fused(DevIndex ix, FusedArgs a)
{
    using M = typename MaskOf<MODE>::T;
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    __shared__ uint32_t s_item;
    const uint32_t ngroups = a.plan[0];
    const uint64_t nd = (uint64_t)ngroups * ix.ntiles;
    const uint64_t ne = a.sp_tot ? (uint64_t)a.sp_tot[1] * a.nslices : 0;
    const uint32_t pg = ngroups ? (blockIdx.x & 7u) % ngroups : 0;
    uint32_t k0 = 0;  // groups, from pg on, that this work-group found exhausted
    for (;;) {
        if (threadIdx.x == 0) {
            uint64_t it = ~0ull;
            while (k0 < ngroups) {
                const uint32_t g = (pg + k0) % ngroups;
                const uint32_t t = atomicAdd(a.counter + 1 + g, 1u);
                if (t < ix.ntiles) {
                    it = (uint64_t)t * ngroups + g;
                    break;
                }
                ++k0;
            }
            if (it == ~0ull) it = nd + atomicAdd(a.counter, 1u);
            s_item = (uint32_t)it;
        }
        __syncthreads();
        const uint64_t item = s_item;
        __syncthreads();
        if (item < nd) {
            const uint32_t tile = (uint32_t)(item / ngroups), g = (uint32_t)(item - (uint64_t)tile * ngroups);
            dense_tile<MODE, NT, true>(ix, tile, reinterpret_cast<const M *>(a.masks) + (uint64_t)g * a.mask_stride,
                                 a.anybms + (uint64_t)g * a.any_stride, a.grp_n[g], a.grp_q + (uint64_t)g * kGroupSlots,
                                 a.out, a.out_stride, a.out_words, a.use_col, lds);
        } else if (item < nd + ne) {
            const uint32_t e = (uint32_t)(item - nd);
            const uint32_t qi = e / a.nslices, sl = e - qi * a.nslices;
            sparse_emit_slice<NT>(ix, a, a.sp_slist[qi], sl, lds);
        } else {
            break;
        }
        __syncthreads();
    }
}

uint32_t fused_lds_bytes(int mode, uint32_t G, uint32_t tile_shift, uint32_t threads, bool use_col)
{
    const uint32_t d = dense_lds_bytes(mode, G, tile_shift, threads, use_col);
    const uint32_t sp = kSpSliceWords * 4;  // one value slice of a sparse query
    return d > sp ? d : sp;
}

template <int MODE, int NT>
static hipError_t occupancy_nt(uint32_t lds, int *per)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per, k_union_fused<MODE, NT>, NT, lds);
}

hipError_t fused_grid(int mode, uint32_t threads, const DevIndex &ix, uint32_t G, bool use_col, uint32_t *grid)
{
    int dev = 0, cus = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const uint32_t lds = fused_lds_bytes(mode, G, ix.tile_shift, threads, use_col);
    if (mode == kModeMask32)
        e = threads == 256 ? occupancy_nt<kModeMask32, 256>(lds, &per) : occupancy_nt<kModeMask32, 512>(lds, &per);
    else if (mode == kModeMask64)
        e = threads == 256 ? occupancy_nt<kModeMask64, 256>(lds, &per) : occupancy_nt<kModeMask64, 512>(lds, &per);
    else if (mode == kModeMask128)
        e = threads == 256 ? occupancy_nt<kModeMask128, 256>(lds, &per) : occupancy_nt<kModeMask128, 512>(lds, &per);
    else
        e = threads == 256 ? occupancy_nt<kModeMask256, 256>(lds, &per) : occupancy_nt<kModeMask256, 512>(lds, &per);
    if (e != hipSuccess) return e;
    *grid = (uint32_t)(cus * (per < 1 ? 1 : per));
    return hipSuccess;
}

hipError_t launch_union_fused(int mode, uint32_t threads, const DevIndex &ix, const FusedArgs &a, uint32_t G,
                              uint32_t grid, hipStream_t s)
{
    launch_fill_zero(a.counter, ((uint64_t)a.ngroups_cap + 1) * 4, s);
    const uint32_t lds = fused_lds_bytes(mode, G, ix.tile_shift, threads, a.use_col != 0);
    if (mode == kModeMask32) {
        if (threads == 256)
            hipLaunchKernelGGL((k_union_fused<kModeMask32, 256>), dim3(grid), dim3(256), lds, s, ix, a);
        else
            hipLaunchKernelGGL((k_union_fused<kModeMask32, 512>), dim3(grid), dim3(512), lds, s, ix, a);
    } else if (mode == kModeMask64) {
        if (threads == 256)
            hipLaunchKernelGGL((k_union_fused<kModeMask64, 256>), dim3(grid), dim3(256), lds, s, ix, a);
        else
            hipLaunchKernelGGL((k_union_fused<kModeMask64, 512>), dim3(grid), dim3(512), lds, s, ix, a);
    } else if (mode == kModeMask128) {
        if (threads == 256)
            hipLaunchKernelGGL((k_union_fused<kModeMask128, 256>), dim3(grid), dim3(256), lds, s, ix, a);
        else
            hipLaunchKernelGGL((k_union_fused<kModeMask128, 512>), dim3(grid), dim3(512), lds, s, ix, a);
    } else {
        if (threads == 256)
            hipLaunchKernelGGL((k_union_fused<kModeMask256, 256>), dim3(grid), dim3(256), lds, s, ix, a);
        else
            hipLaunchKernelGGL((k_union_fused<kModeMask256, 512>), dim3(grid), dim3(512), lds, s, ix, a);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bitmap -> ascending ids (CollectIdsFromBitmap, test/adsfilter_q_all.cu:108-119)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_popcount(const uint32_t *__restrict__ bm, uint64_t stride, uint64_t words,
                                                  uint32_t nchunk, uint32_t *ccnt, const uint8_t *skip)
{
    __shared__ uint32_t sh[4];
    const uint32_t q = blockIdx.y, c = blockIdx.x;
    if (skip && skip[q] == kSpSlice) return;
    const uint64_t w0 = (uint64_t)c * kCompactWords;
    const uint64_t w1 = w0 + kCompactWords < words ? w0 + kCompactWords : words;
    const uint32_t *src = bm + (uint64_t)q * stride;
    uint32_t s = 0;
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        // 16-byte loads over the chunk's whole quads (w0 is a multiple of 4),
        // then the words past the last one
        const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
        const uint64_t e4 = w1 / 4;
        for (uint64_t i = w0 / 4 + threadIdx.x; i < e4; i += 256) {
            const uint4 v = s4[i];
            s += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
        }
        for (uint64_t i = e4 * 4 + threadIdx.x; i < w1; i += 256) s += __popc(src[i]);
    } else {
        for (uint64_t i = w0 + threadIdx.x; i < w1; i += 256) s += __popc(src[i]);
    }
    s = wave_sum(s);
    if (lane_id() == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) ccnt[(uint64_t)q * nchunk + c] = sh[0] + sh[1] + sh[2] + sh[3];
}

void launch_popcount(const uint32_t *bm, uint64_t stride, uint64_t words, uint32_t nq, uint32_t *ccnt, hipStream_t s,
                     const uint8_t *skip)
{
    if (!nq || !words) return;
    uint32_t nchunk = (uint32_t)((words + kCompactWords - 1) / kCompactWords);
    for (uint32_t q0 = 0; q0 < nq; q0 += kMaxGridY) {
        const uint32_t n = nq - q0 < kMaxGridY ? nq - q0 : kMaxGridY;
        hipLaunchKernelGGL(k_popcount, dim3(nchunk, n), dim3(256), 0, s, bm + (uint64_t)q0 * stride, stride, words,
                           nchunk, ccnt + (uint64_t)q0 * nchunk, skip ? skip + q0 : nullptr);
    }
}

// A work-group expands one chunk of 2048 words in 8 steps of 256 words; in
// step i, wave w takes words 256 i + 64 w .. + 63 (one per lane, all eight
// kept in registers from the counting pass), so the four waves of a step
// write adjacent pieces of the output (6% faster than each wave taking its
// own 512-word range).  Per 64-word piece a wave prefix sum of popcounts
// places every lane's ids in a wave-private LDS staging area, shifted by the
// output offset modulo 4 so that staging quads and 16-byte output quads line
// up; the wave then stores whole quads with 128-bit writes, two LDS reads in
// flight before their stores (ALIGNED: ids + q * ids_stride is 16-byte
// aligned), and lanes 0 and 1 the partial end quads dword by dword.  The
// popcount prefix is a DPP scan (no LDS round trips): with both, 2% less
// writer time.
constexpr uint32_t kStage = 64 * 32 + 4;
// Per query, in place: chunk counts -> exclusive prefix; counts[q] = total.
__global__ __launch_bounds__(kScanNT) void k_chunk_scan(uint32_t *ccnt, uint32_t nchunk, uint32_t *counts)
{
    __shared__ unsigned long long sh[kScanNT / 64];
    uint32_t *cc = ccnt + (uint64_t)blockIdx.x * nchunk;
    unsigned long long carry = 0;
    for (uint32_t base = 0; base < nchunk; base += kScanNT) {
        const uint32_t j = base + threadIdx.x;
        const unsigned long long x = j < nchunk ? cc[j] : 0;
        unsigned long long tot;
        const unsigned long long e = block_excl_scan(x, sh, &tot);
        if (j < nchunk) cc[j] = (uint32_t)(carry + e);
        carry += tot;
    }
    if (threadIdx.x == 0 && counts) counts[blockIdx.x] = (uint32_t)carry;
}

// One wave's 64-word piece of a bitmap: lane l holds word v, whose first
// value is base; the piece's ids go to dst[off ...] in ascending order,
// through the wave's staging area stg (kStage words).  Returns nothing; the
// caller advances off by the piece's popcount.
template <bool ALIGNED>
__device__ __forceinline__ void expand_piece(uint32_t v, uint32_t base, uint64_t off, uint32_t *dst, uint32_t *stg,
                                             uint32_t lane)
{
    const uint32_t cnt = __popc(v);
    const uint32_t incl = wave_incl_scan_dpp(cnt);
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    if (tot == 0) return;
    const uint32_t sh = ALIGNED ? (uint32_t)(off & 3) : 0u;
    uint32_t p = sh + incl - cnt;
    while (v) {
        stg[p++] = base + __builtin_ctz(v);
        v &= v - 1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (ALIGNED) {
        // staging index j <-> output index off - sh + j
        uint32_t *d0 = dst + (off - sh);
        const uint32_t end = sh + tot;
        // whole quads [k0, k1); the partial first (sh > 0) and last (end % 4) quads by lanes 0 and 1
        const uint32_t k0 = (sh + 3) >> 2, k1 = end >> 2;
        if (lane == 0 && sh) {
            for (uint32_t j = sh; j < 4 && j < end; ++j) d0[j] = stg[j];
        }
        if (lane == 1 && (end & 3) && (end >> 2) >= k0) {
            for (uint32_t j = end & ~3u; j < end; ++j) d0[j] = stg[j];
        }
        uint32_t k = k0 + lane;
        for (; k + 64 < k1; k += 128) {
            const uint4 a = *reinterpret_cast<const uint4 *>(stg + 4 * k);
            const uint4 b = *reinterpret_cast<const uint4 *>(stg + 4 * (k + 64));
            *reinterpret_cast<uint4 *>(d0 + 4 * k) = a;
            *reinterpret_cast<uint4 *>(d0 + 4 * (k + 64)) = b;
        }
        if (k < k1) *reinterpret_cast<uint4 *>(d0 + 4 * k) = *reinterpret_cast<const uint4 *>(stg + 4 * k);
    } else {
        for (uint32_t j = lane; j < tot; j += 64) dst[off + j] = stg[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ccnt: per-chunk exclusive prefixes (k_chunk_scan).
template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_write_ids(const uint32_t *__restrict__ bm, uint64_t stride, uint64_t words,
                                                   uint32_t nchunk, const uint32_t *__restrict__ ccnt,
                                                   uint32_t *__restrict__ ids, uint64_t ids_stride,
                                                   const uint64_t *__restrict__ id_off)
{
    __shared__ __attribute__((aligned(16))) uint32_t stage[4][kStage];
    __shared__ uint32_t wcnt[8][4];
    const uint32_t q = blockIdx.y, c = blockIdx.x;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t prefix = ccnt[(uint64_t)q * nchunk + c];  // ids of the query's earlier chunks
    const uint32_t *src = bm + (uint64_t)q * stride;
    // step i of wave w covers words c*2048 + (4i + w)*64 ...: the four waves
    // of a step write adjacent pieces of the output
    const uint64_t ww0 = (uint64_t)c * kCompactWords + (uint64_t)wave * 64;
    uint32_t x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t wi = ww0 + i * 256 + lane;
        x[i] = wi < words ? src[wi] : 0u;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t n = wave_sum((uint32_t)__popc(x[i]));
        if (lane == 0) wcnt[i][wave] = n;
    }
    __syncthreads();
    uint64_t run = prefix;
    uint32_t *dst = ids + (id_off ? id_off[q] : (uint64_t)q * ids_stride);
    uint32_t *stg = stage[wave];
#pragma unroll 1
    for (int i = 0; i < 8; ++i) {
        uint64_t off = run;
        for (uint32_t w = 0; w < wave; ++w) off += wcnt[i][w];
        run += wcnt[i][0] + wcnt[i][1] + wcnt[i][2] + wcnt[i][3];
        uint32_t v = x[0];
#pragma unroll
        for (int k = 1; k < 8; ++k) v = (i == k) ? x[k] : v;
        expand_piece<ALIGNED>(v, (uint32_t)((ww0 + (uint64_t)i * 256 + lane) * 32), off, dst, stg, lane);
    }
}

void launch_write_ids(const uint32_t *bm, uint64_t stride, uint64_t words, uint32_t nq, uint32_t *ccnt,
                      uint32_t *ids, uint64_t ids_stride, const uint64_t *id_off, bool off_aligned,
                      uint32_t *counts, hipStream_t s)
{
    if (!nq || !words) return;
    uint32_t nchunk = (uint32_t)((words + kCompactWords - 1) / kCompactWords);
    hipLaunchKernelGGL(k_chunk_scan, dim3(nq), dim3(kScanNT), 0, s, ccnt, nchunk, counts);
    if (!ids) return;
    const bool aligned = ((uintptr_t)ids % 16 == 0) && (id_off ? off_aligned : ids_stride % 4 == 0);
    for (uint32_t q0 = 0; q0 < nq; q0 += kMaxGridY) {
        const uint32_t n = nq - q0 < kMaxGridY ? nq - q0 : kMaxGridY;
        const uint32_t *b = bm + (uint64_t)q0 * stride;
        const uint32_t *cc = ccnt + (uint64_t)q0 * nchunk;
        uint32_t *dst = id_off ? ids : ids + (uint64_t)q0 * ids_stride;
        const uint64_t *off = id_off ? id_off + q0 : nullptr;
        if (aligned)
            hipLaunchKernelGGL(k_write_ids<true>, dim3(nchunk, n), dim3(256), 0, s, b, stride, words, nchunk, cc, dst,
                               ids_stride, off);
        else
            hipLaunchKernelGGL(k_write_ids<false>, dim3(nchunk, n), dim3(256), 0, s, b, stride, words, nchunk, cc,
                               dst, ids_stride, off);
    }
}

// ---------------------------------------------------------------------------
// id-only batches by value slice: each query's routed records OR-ed into an
// LDS slice bitmap, counted, and expanded straight to ids (no bitmap in HBM,
// no sort)
// ---------------------------------------------------------------------------
// Status word of a (query, slice) item in the one-pass prefix: kLbInc |
// inclusive prefix, kLbAgg | the item's own count, or 0 (not yet published);
// read and written with relaxed agent-scope atomics (each carries its own
// value, so no other memory is ordered by it).
constexpr unsigned long long kLbAgg = 1ull << 62, kLbInc = 1ull << 63, kLbVal = kLbAgg - 1;
// A look-back that polls this often gives up and sets the error word (a
// guard against a hang; a correct launch never comes near it).
constexpr uint32_t kLbSpinMax = 1u << 22;

__device__ __forceinline__ unsigned long long lb_load(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of item j0 (> 0) among the items whose status words start
// at sq, by one wave: lane l reads item j - 1 - l, and each window of 64 is
// summed up to its first inclusive prefix.  Items are taken in ticket order,
// so every item waited on belongs to a running work-group that publishes its
// count before it waits itself.
__device__ __forceinline__ unsigned long long lb_prefix(unsigned long long *sq, uint32_t j0, uint32_t *err)
{
    const uint32_t lane = lane_id();
    unsigned long long excl = 0;
    uint32_t j = j0, spins = 0;
    while (j > 0) {
        const bool valid = lane < j;
        const uint32_t idx = j - 1 - lane;
        // lanes before item 0 never count: item 0 always holds an inclusive prefix
        unsigned long long v = valid ? lb_load(&sq[idx]) : kLbInc;
        for (;;) {
            const unsigned long long incm = __ballot((v & kLbInc) != 0);
            const uint32_t first = incm ? (uint32_t)__builtin_ctzll(incm) : 64u;
            const unsigned long long need = first >= 63 ? ~0ull : ((2ull << first) - 1);
            if (!(__ballot(v == 0) & need)) {
                excl += wave_sum64(lane <= first ? (v & kLbVal) : 0ull);
                j = first < 64 ? 0 : j - 64;
                break;
            }
            if (++spins > kLbSpinMax) {
                if (lane == 0) atomicOr(err, 1u);
                return excl;
            }
            if (v == 0) v = lb_load(&sq[idx]);
        }
    }
    return excl;
}

// Persistent: items (query qi, slice s) in ticket order.  ORs the slice's
// routed records into LDS, counts them in steps of 256 words (wave w takes
// words 256 i + 64 w ..), publishes the count, finds the query's ids before
// the slice by look-back, then expands each 64-word piece to ids.  The last
// slice of a query writes its count.
template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_ids_slices(DevIndex ix, IdsSliceArgs a)
{
    constexpr uint32_t NT = 256, NS = kSpSliceWords / 256;
    __shared__ __attribute__((aligned(16))) uint32_t slice[kSpSliceWords];
    __shared__ __attribute__((aligned(16))) uint32_t stage[4][kStage];
    __shared__ uint32_t wcnt[NS][4];
    __shared__ uint32_t s_item, s_any;
    __shared__ unsigned long long s_excl;
    __shared__ uint32_t s_lo[NT], s_pre[NT];
    __shared__ unsigned long long s_sh[NT / 64];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t S = a.nslices;
    const uint64_t nitems = (uint64_t)a.nq * S;
    for (;;) {
        if (threadIdx.x == 0) {
            s_item = atomicAdd(a.counter, 1u);
            s_any = 0;
        }
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= nitems) break;
        const uint32_t qi = item / S, s = item - qi * S;
        const uint32_t w0 = s * kSpSliceWords;
        const uint32_t nch = a.cpre[qi + 1] - a.cpre[qi];
        const uint16_t *co = a.choff + (uint64_t)a.cpre[qi] * (S + 1);
        const uint32_t *refs = a.refs + a.rbase[qi];
        // does any chunk hold records of this slice?
        for (uint32_t j = threadIdx.x; j < nch; j += NT)
            if (co[(uint64_t)j * (S + 1) + s + 1] != co[(uint64_t)j * (S + 1) + s]) s_any = 1;
        for (uint32_t i = threadIdx.x; i < kSpSliceWords / 4; i += NT)
            reinterpret_cast<uint4 *>(slice)[i] = make_uint4(0, 0, 0, 0);
        __syncthreads();
        const bool any = s_any != 0;
        if (any) {
            or_slice_refs<NT>(ix, co, refs, nch, S, a.chunk, s, w0, slice, s_lo, s_pre, s_sh);
#pragma unroll 1
            for (uint32_t i = 0; i < NS; ++i) {
                const uint32_t n = (uint32_t)__builtin_amdgcn_readlane(
                    (int)wave_incl_scan_dpp((uint32_t)__popc(slice[i * 256 + wave * 64 + lane])), 63);
                if (lane == 0) wcnt[i][wave] = n;
            }
        } else if (threadIdx.x < 4 * NS) {
            wcnt[threadIdx.x >> 2][threadIdx.x & 3] = 0;
        }
        __syncthreads();
        if (wave == 0) {
            uint32_t mine = 0;
            for (uint32_t i = lane; i < 4 * NS; i += 64) mine += wcnt[i >> 2][i & 3];
            const uint32_t agg = wave_sum(mine);
            unsigned long long *sq = a.st + (uint64_t)qi * S;
            if (lane == 0) lb_store(&sq[s], (s == 0 ? kLbInc : kLbAgg) | agg);
            const unsigned long long excl = s ? lb_prefix(sq, s, a.err) : 0ull;
            if (lane == 0) {
                if (s) lb_store(&sq[s], kLbInc | (excl + agg));
                if (s + 1 == S) a.counts[a.q0 + qi] = (uint32_t)(excl + agg);
                s_excl = excl;
            }
        }
        __syncthreads();
        if (any) {
            uint64_t run = s_excl;
            uint32_t *dst = a.ids + a.id_off[a.q0 + qi];
            uint32_t *stg = stage[wave];
#pragma unroll 1
            for (uint32_t i = 0; i < NS; ++i) {
                uint64_t off = run;
                for (uint32_t w = 0; w < wave; ++w) off += wcnt[i][w];
                run += wcnt[i][0] + wcnt[i][1] + wcnt[i][2] + wcnt[i][3];
                const uint32_t word = i * 256 + wave * 64 + lane;
                expand_piece<ALIGNED>(slice[word], (w0 + word) * 32, off, dst, stg, lane);
            }
        }
        __syncthreads();
    }
}

uint32_t ids_slices_grid()
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev))
        return 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ids_slices<true>, 256, 0) != hipSuccess || per < 1)
        per = 1;
    return (uint32_t)(cus * per);
}

hipError_t launch_ids_slices(const DevIndex &ix, const IdsSliceArgs &a, bool aligned, uint32_t grid, hipStream_t s)
{
    if (aligned)
        hipLaunchKernelGGL(k_ids_slices<true>, dim3(grid), dim3(256), 0, s, ix, a);
    else
        hipLaunchKernelGGL(k_ids_slices<false>, dim3(grid), dim3(256), 0, s, ix, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// cost-model features and the threshold rule
// ---------------------------------------------------------------------------
// One 256-thread work-group per query; every thread keeps four keys' row
// loads in flight.  All sums are integers (exact in any order), so the
// features equal the host's bit for bit up to the final float conversion.
constexpr uint32_t kFeatNT = 256;
__global__ __launch_bounds__(kFeatNT) void k_features(DevIndex ix, const uint32_t *__restrict__ keys,
                                                      const uint64_t *__restrict__ qoff, uint64_t R, float *feats,
                                                      uint64_t *Mout)
{
    constexpr uint32_t kBytes[6] = {4, 8, 12, 20, 36, 36};
    constexpr uint32_t U = 4;
    __shared__ unsigned long long sh[kFeatNT / 64][8];
    const uint32_t q = blockIdx.x, lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t a = qoff[q], e = qoff[q + 1];
    unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // valid M sq mx bytes nv a1 bs
    for (uint64_t j0 = a + threadIdx.x; j0 < e; j0 += (uint64_t)kFeatNT * U) {
        uint32_t k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint64_t j = j0 + (uint64_t)u * kFeatNT;
            k[u] = j < e ? keys[j] : 0xffffffffu;
        }
        uint4 lo0[U], lo1[U], hi0[U], hi1[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (k[u] < ix.nkeys) {
                const uint4 *r = reinterpret_cast<const uint4 *>(ix.seg8 + (uint64_t)k[u] * 8);
                lo0[u] = r[0];
                lo1[u] = r[1];
                hi0[u] = r[2];
                hi1[u] = r[3];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (k[u] >= ix.nkeys) continue;
            const uint32_t c[8] = {hi0[u].x - lo0[u].x, hi0[u].y - lo0[u].y, hi0[u].z - lo0[u].z,
                                   hi0[u].w - lo0[u].w, hi1[u].x - lo1[u].x, hi1[u].y - lo1[u].y,
                                   hi1[u].z - lo1[u].z, hi1[u].w - lo1[u].w};
            const unsigned long long tot = c[6];
            v[0] += 1;
            v[1] += tot;
            v[2] += tot * tot;
            v[3] = tot > v[3] ? tot : v[3];
#pragma unroll
            for (int t = 0; t < 6; ++t) v[4] += (unsigned long long)c[t] * kBytes[t];
            v[5] += c[7];
            v[6] += c[0];
            v[7] += c[5];
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (i == 3) {
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) {
                const unsigned long long o = __shfl_xor(v[3], d, 64);
                v[3] = o > v[3] ? o : v[3];
            }
        } else {
            v[i] = wave_sum64(v[i]);
        }
    }
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < 8; ++i) sh[wave][i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t w = 0; w < kFeatNT / 64; ++w)
            for (int i = 0; i < 8; ++i) t[i] = i == 3 ? (sh[w][i] > t[i] ? sh[w][i] : t[i]) : t[i] + sh[w][i];
        const unsigned long long valid = t[0], M = t[1], mx = t[3], bytes = t[4], nv = t[5], a1 = t[6], bs = t[7];
        const double n = (double)(e - a);
        const double dv = valid ? (double)valid : 1.0;
        const double mean = (double)M / dv;
        double var = (double)t[2] / dv - mean * mean;
        if (var < 0) var = 0;
        const double dM = M ? (double)M : 1.0;
        const double nvd = ix.nvalues ? (double)ix.nvalues : 1.0;
        float *f = feats + (uint64_t)q * AB_NFEATURES;
        f[0] = (float)log1p(n);
        f[1] = (float)log1p((double)valid);
        f[2] = (float)log1p((double)M);
        f[3] = (float)log1p((double)mx);
        f[4] = (float)log1p(sqrt(var));
        f[5] = (float)log1p((double)bytes);
        f[6] = (float)((double)M / (R ? (double)R : 1.0));
        const double hd = (double)nv / nvd;
        f[7] = (float)(hd < 1.0 ? hd : 1.0);
        f[8] = (float)(1.0 - exp(-hd));
        f[9] = (float)((double)valid / (ix.nkeys ? (double)ix.nkeys : 1.0));
        f[10] = (float)((double)a1 / dM);
        f[11] = (float)((double)bs / dM);
        if (Mout) Mout[q] = M;
    }
}

void launch_features(const DevIndex &ix, const uint32_t *keys, const uint64_t *qoff, uint32_t nq, uint64_t R,
                     float *feats, uint64_t *M, hipStream_t s)
{
    if (!nq) return;
    hipLaunchKernelGGL(k_features, dim3(nq), dim3(kFeatNT), 0, s, ix, keys, qoff, R, feats, M);
}

__global__ void k_decide_threshold(const uint64_t *M, uint32_t nq, double thr, uint8_t *sel)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) sel[q] = (double)M[q] >= thr ? (uint8_t)AB_PATH_DENSE : (uint8_t)AB_PATH_SPARSE;
}

void launch_decide_threshold(const uint64_t *M, uint32_t nq, double thr, uint8_t *sel, hipStream_t s)
{
    if (!nq) return;
    hipLaunchKernelGGL(k_decide_threshold, dim3((nq + 255) / 256), dim3(256), 0, s, M, nq, thr, sel);
}

}  // namespace ab
