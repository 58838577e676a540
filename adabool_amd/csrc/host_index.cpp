// host_index.cpp -- host-side construction of the AdaBool compressed index.
//
// Semantics follow ivf_list::BuildFrom / GroupKeys
// (src/AdaBool/inverted_list.h:348-470): values of a key are grouped by
// 256-value window w = v & 0xffffff00; a window holding n low bytes becomes one
// record of type ARRAY1 (n<=1), ARRAY4 (<=5), ARRAY8 (<=9), ARRAY16 (<=17),
// ARRAY32 (<=33) or BITSET, with header w | low[0] (BITSET: w) and payload
// low[1..] zero-padded (BITSET: a 256-bit set).  Records of a type are ordered
// by (key, w).  Construction itself is a parallel radix sort of packed pairs
// followed by a linear scan, not the reference's std::map-of-maps.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

#include "internal.h"

namespace ab {

int default_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 4;
    if (n > 16) n = 16;
    return (int)n;
}

uint64_t stride_bytes(uint32_t nvalues)
{
    uint64_t b = ((uint64_t)nvalues + 7) / 8;
    return (b + 15) / 16 * 16;
}

uint32_t auto_tile_shift(uint32_t nvalues)
{
    // Aim for >= 1024 value tiles (work-groups) with tiles of 2^8..2^13 values.
    uint32_t s = 8;
    while (s < 13 && ((uint64_t)nvalues >> (s + 1)) >= 1024) ++s;
    return s;
}

static int bits_for(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

// LSD radix sort of 64-bit words on their low `bits` bits, 11-bit digits,
// per-thread histograms.
static void radix_sort_u64(std::vector<uint64_t> &a, int bits, int nthreads)
{
    const int kDigit = 11, kRadix = 1 << kDigit;
    uint64_t n = a.size();
    if (n < 2 || bits <= 0) return;
    std::vector<uint64_t> b(n);
    std::vector<uint64_t> hist((size_t)nthreads * kRadix);
    for (int shift = 0; shift < bits; shift += kDigit) {
        std::fill(hist.begin(), hist.end(), 0);
        parallel_ranges(n, nthreads, [&](uint64_t lo, uint64_t hi, int t) {
            uint64_t *h = &hist[(size_t)t * kRadix];
            for (uint64_t i = lo; i < hi; ++i) h[(a[i] >> shift) & (kRadix - 1)]++;
        });
        uint64_t run = 0;
        for (int d = 0; d < kRadix; ++d)
            for (int t = 0; t < nthreads; ++t) {
                uint64_t c = hist[(size_t)t * kRadix + d];
                hist[(size_t)t * kRadix + d] = run;
                run += c;
            }
        parallel_ranges(n, nthreads, [&](uint64_t lo, uint64_t hi, int t) {
            uint64_t *h = &hist[(size_t)t * kRadix];
            for (uint64_t i = lo; i < hi; ++i) b[h[(a[i] >> shift) & (kRadix - 1)]++] = a[i];
        });
        a.swap(b);
    }
}

static inline int bucket_type(uint64_t n)
{
    if (n <= 1) return AB_ARRAY1;
    if (n <= 5) return AB_ARRAY4;
    if (n <= 9) return AB_ARRAY8;
    if (n <= 17) return AB_ARRAY16;
    if (n <= 33) return AB_ARRAY32;
    return AB_BITSET;
}

ab_status build_host_index(const uint32_t *keys, const uint32_t *vals, uint64_t nnz,
                           uint32_t nkeys, uint32_t nvalues, uint32_t tile_shift,
                           int nthreads, HostIndex *out)
{
    if (nnz && (!keys || !vals)) {
        set_error("ab_index_build: NULL pair arrays");
        return AB_INVALID_ARG;
    }
    if (tile_shift != 0 && (tile_shift < 8 || tile_shift > 16)) {
        set_error("ab_index_build: tile_shift must be 0 or 8..16");
        return AB_INVALID_ARG;
    }
    if (nthreads <= 0) nthreads = default_threads();
    // ---- key / value space --------------------------------------------------
    std::vector<uint32_t> mk(nthreads, 0), mv(nthreads, 0);
    parallel_ranges(nnz, nthreads, [&](uint64_t lo, uint64_t hi, int t) {
        uint32_t a = 0, b = 0;
        for (uint64_t i = lo; i < hi; ++i) {
            a = std::max(a, keys[i]);
            b = std::max(b, vals[i]);
        }
        mk[t] = a;
        mv[t] = b;
    });
    uint64_t need_k = 0, need_v = 0;
    if (nnz) {
        need_k = (uint64_t)*std::max_element(mk.begin(), mk.end()) + 1;
        need_v = (uint64_t)*std::max_element(mv.begin(), mv.end()) + 1;
    }
    if (need_k > 0xffffffffull || need_v > 0xffffffffull) {
        set_error("ab_index_build: key or value 0xffffffff is not representable (space would be 2^32)");
        return AB_UNSUPPORTED;
    }
    if (nkeys == 0) nkeys = (uint32_t)need_k;
    if (nvalues == 0) nvalues = (uint32_t)need_v;
    if (nkeys < need_k || nvalues < need_v) {
        set_error("ab_index_build: nkeys/nvalues smaller than max+1 of the pairs");
        return AB_INVALID_ARG;
    }
    HostIndex &ix = *out;
    ix = HostIndex();
    ix.nkeys = nkeys;
    ix.nvalues = nvalues;
    ix.tile_shift = tile_shift ? tile_shift : auto_tile_shift(nvalues);

    // ---- sort packed (key, value) pairs and drop duplicates -----------------
    int vbits = bits_for(nvalues ? nvalues - 1 : 0);
    int kbits = bits_for(nkeys ? nkeys - 1 : 0);
    if (vbits + kbits > 64) {
        set_error("ab_index_build: key and value spaces too large to pack");
        return AB_UNSUPPORTED;
    }
    std::vector<uint64_t> kv(nnz);
    parallel_ranges(nnz, nthreads, [&](uint64_t lo, uint64_t hi, int) {
        for (uint64_t i = lo; i < hi; ++i) kv[i] = ((uint64_t)keys[i] << vbits) | vals[i];
    });
    radix_sort_u64(kv, vbits + kbits, nthreads);
    if (!kv.empty()) {
        uint64_t w = 1;
        for (uint64_t i = 1; i < kv.size(); ++i)
            if (kv[i] != kv[w - 1]) kv[w++] = kv[i];
        kv.resize(w);
    }
    const uint64_t vmask = vbits ? ((vbits == 64) ? ~0ull : ((1ull << vbits) - 1)) : 0;
    auto K = [&](uint64_t i) { return (uint32_t)(vbits ? (kv[i] >> vbits) : kv[i]); };
    auto V = [&](uint64_t i) { return (uint32_t)(kv[i] & vmask); };
    ix.nnz = kv.size();
    if (ix.nnz >= 0xffffffffull) {
        set_error("ab_index_build: more than 2^32-1 distinct pairs");
        return AB_UNSUPPORTED;
    }

    // ---- split the sorted pairs at key boundaries for parallel scans -------
    const uint64_t n = kv.size();
    std::vector<uint64_t> cut(nthreads + 1, n);
    cut[0] = 0;
    for (int t = 1; t < nthreads; ++t) {
        uint64_t p = n * t / nthreads;
        if (p < cut[t - 1]) p = cut[t - 1];
        while (p < n && p > 0 && K(p) == K(p - 1)) ++p;
        cut[t] = p;
    }
    // pass 1: per-key counts into seg8 rows k+1 and per-chunk record counts
    ix.seg8.assign(((uint64_t)nkeys + 1) * kSegStride, 0);
    std::vector<uint64_t> chunk_cnt((size_t)nthreads * kTypes, 0);
    parallel_ranges(nthreads, nthreads, [&](uint64_t c0, uint64_t c1, int) {
        for (uint64_t c = c0; c < c1; ++c) {
            uint64_t i = cut[c], e = cut[c + 1];
            while (i < e) {
                uint32_t key = K(i);
                uint32_t *row = &ix.seg8[((uint64_t)key + 1) * kSegStride];
                while (i < e && K(i) == key) {
                    uint32_t win = V(i) & 0xffffff00u;
                    uint64_t j = i;
                    while (j < e && K(j) == key && (V(j) & 0xffffff00u) == win) ++j;
                    int t = bucket_type(j - i);
                    row[t]++;
                    row[6]++;
                    row[7] += (uint32_t)(j - i);
                    chunk_cnt[c * kTypes + t]++;
                    i = j;
                }
            }
        }
    });
    // record totals must fit the u32 prefix arrays
    uint64_t tot[kTypes] = {0};
    for (int c = 0; c < nthreads; ++c)
        for (int t = 0; t < kTypes; ++t) tot[t] += chunk_cnt[c * kTypes + t];
    uint64_t R = 0;
    for (int t = 0; t < kTypes; ++t) R += tot[t];
    if (R >= 0xffffffffull) {
        set_error("ab_index_build: more than 2^32-1 records");
        return AB_UNSUPPORTED;
    }
    // prefix over keys per slot
    for (uint64_t k = 0; k < nkeys; ++k) {
        uint32_t *a = &ix.seg8[k * kSegStride], *b = a + kSegStride;
        for (int s = 0; s < kSegStride; ++s) b[s] += a[s];
    }
    // chunk bases per type
    std::vector<uint64_t> chunk_base((size_t)nthreads * kTypes, 0);
    for (int t = 0; t < kTypes; ++t) {
        uint64_t run = 0;
        for (int c = 0; c < nthreads; ++c) {
            chunk_base[c * kTypes + t] = run;
            run += chunk_cnt[c * kTypes + t];
        }
    }
    std::vector<uint32_t> rkey[kTypes];  // key of each key-major record
    for (int t = 0; t < kTypes; ++t) {
        ix.hdr[t].resize(tot[t]);
        ix.pay[t].assign(tot[t] * kPayloadBytes[t], 0);
        rkey[t].resize(tot[t]);
    }
    // pass 2: fill headers and payloads (GroupKeys, inverted_list.h:390-470)
    parallel_ranges(nthreads, nthreads, [&](uint64_t c0, uint64_t c1, int) {
        for (uint64_t c = c0; c < c1; ++c) {
            uint64_t pos[kTypes];
            for (int t = 0; t < kTypes; ++t) pos[t] = chunk_base[c * kTypes + t];
            uint64_t i = cut[c], e = cut[c + 1];
            while (i < e) {
                uint32_t key = K(i);
                uint32_t win = V(i) & 0xffffff00u;
                uint64_t j = i;
                while (j < e && K(j) == key && (V(j) & 0xffffff00u) == win) ++j;
                uint64_t cnt = j - i;
                int t = bucket_type(cnt);
                uint64_t r = pos[t]++;
                rkey[t][r] = key;
                if (t == AB_BITSET) {
                    ix.hdr[t][r] = win;
                    uint8_t *p = &ix.pay[t][r * 32];
                    for (uint64_t b = i; b < j; ++b) {
                        uint32_t lo = V(b) & 0xff;
                        p[lo >> 3] |= (uint8_t)(1u << (lo & 7));
                    }
                } else {
                    ix.hdr[t][r] = win | (V(i) & 0xff);
                    uint8_t *p = ix.pay[t].data() + r * kPayloadBytes[t];
                    for (uint64_t b = 1; b < cnt; ++b) p[b - 1] = (uint8_t)(V(i + b) & 0xff);
                }
                i = j;
            }
        }
    });
    std::vector<uint64_t>().swap(kv);

    // ---- value-tiled layout ----------------------------------------------------
    const uint32_t ts = ix.tile_shift;
    const uint32_t ntiles = (uint32_t)(((uint64_t)nvalues + (1ull << ts) - 1) >> ts);
    ix.ntiles = ntiles;
    // per type: stable counting sort of records by tile
    std::vector<std::vector<uint32_t>> tstart(kTypes);  // ntiles+1 per type
    std::vector<std::vector<uint32_t>> vkey(kTypes);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < kTypes; ++t)
            th.emplace_back([&, t]() {
                const uint64_t nr = ix.hdr[t].size();
                const int pb = kPayloadBytes[t];
                std::vector<uint32_t> &st = tstart[t];
                st.assign((size_t)ntiles + 1, 0);
                for (uint64_t r = 0; r < nr; ++r) st[(ix.hdr[t][r] >> ts) + 1]++;
                for (uint32_t j = 0; j < ntiles; ++j) st[j + 1] += st[j];
                std::vector<uint32_t> cur(st.begin(), st.end() - 1);
                ix.vhdr[t].resize(nr);
                ix.vpay[t].resize(nr * pb);
                vkey[t].resize(nr);
                for (uint64_t r = 0; r < nr; ++r) {
                    uint32_t d = cur[ix.hdr[t][r] >> ts]++;
                    ix.vhdr[t][d] = ix.hdr[t][r];
                    vkey[t][d] = rkey[t][r];
                    if (pb) memcpy(&ix.vpay[t][(uint64_t)d * pb], &ix.pay[t][r * pb], pb);
                }
            });
        for (auto &x : th) x.join();
    }
    for (int t = 0; t < kTypes; ++t) std::vector<uint32_t>().swap(rkey[t]);
    // blocks: per tile, per type, ceil(count/1024) blocks.  Two passes.
    auto width_of = [](uint32_t span) -> uint32_t { return span < 256u ? 1 : (span < 65536u ? 2 : 4); };
    std::vector<uint64_t> tile_nblk(ntiles + 1, 0), tile_nbytes(ntiles + 1, 0);
    parallel_ranges(ntiles, nthreads, [&](uint64_t j0, uint64_t j1, int) {
        for (uint64_t j = j0; j < j1; ++j) {
            uint64_t nb = 0, bytes = 0;
            for (int t = 0; t < kTypes; ++t) {
                uint32_t s = tstart[t][j], e = tstart[t][j + 1];
                for (uint32_t b = s; b < e; b += kBlockRecs) {
                    uint32_t nxt = (b + kBlockRecs < e) ? vkey[t][b + kBlockRecs] : vkey[t][e - 1];
                    bytes += (uint64_t)kBlockRecs * width_of(nxt - vkey[t][b]);
                    ++nb;
                }
            }
            tile_nblk[j + 1] = nb;
            tile_nbytes[j + 1] = bytes;
        }
    });
    for (uint32_t j = 0; j < ntiles; ++j) {
        tile_nblk[j + 1] += tile_nblk[j];
        tile_nbytes[j + 1] += tile_nbytes[j];
    }
    if ((tile_nbytes[ntiles] >> 10) >= 0xffffffffull) {
        set_error("ab_index_build: key stream exceeds 4 TiB");
        return AB_UNSUPPORTED;
    }
    ix.tile_blk.resize(ntiles + 1);
    for (uint32_t j = 0; j <= ntiles; ++j) ix.tile_blk[j] = (uint32_t)tile_nblk[j];
    ix.vblk.resize(tile_nblk[ntiles]);
    ix.vkeys.assign(tile_nbytes[ntiles] + 64, 0);  // +64: loads never run past the end
    parallel_ranges(ntiles, nthreads, [&](uint64_t j0, uint64_t j1, int) {
        for (uint64_t j = j0; j < j1; ++j) {
            uint64_t bi = tile_nblk[j], off = tile_nbytes[j];
            for (int t = 0; t < kTypes; ++t) {
                uint32_t s = tstart[t][j], e = tstart[t][j + 1];
                for (uint32_t b = s; b < e; b += kBlockRecs) {
                    uint32_t nv = std::min<uint32_t>(kBlockRecs, e - b);
                    uint32_t k0 = vkey[t][b];
                    uint32_t nxt = (b + kBlockRecs < e) ? vkey[t][b + kBlockRecs] : vkey[t][e - 1];
                    uint32_t w = width_of(nxt - k0);
                    uint32_t wl = w == 1 ? 0 : (w == 2 ? 1 : 2);
                    BlockMeta m;
                    m.start_key = k0;
                    m.off_kb = (uint32_t)(off >> 10);
                    m.tw = (uint32_t)t | (wl << 4) | (nv << 8);
                    m.rec_base = b;
                    ix.vblk[bi++] = m;
                    uint8_t *dst = &ix.vkeys[off];
                    for (uint32_t r = 0; r < nv; ++r) {
                        uint32_t d = vkey[t][b + r] - k0;
                        uint32_t lane = r & 63, slot = r >> 6;
                        uint8_t *p = dst + ((uint64_t)lane * kSlots + slot) * w;
                        for (uint32_t q = 0; q < w; ++q) p[q] = (uint8_t)(d >> (8 * q));
                    }
                    off += (uint64_t)kBlockRecs * w;
                }
            }
        }
    });
    return AB_OK;
}

// Values of one record (header h, payload p of type t) in emit order: the
// header value, then payload bytes [0, kMinNum) always and later bytes up to
// the first zero; a bitset's set bits.  f(value) per value.
template <class F>
static void for_record_values(int t, uint32_t h, const uint8_t *p, F &&f)
{
    if (t == AB_BITSET) {
        for (int x = 0; x < 256; ++x)
            if (p[x >> 3] & (1u << (x & 7))) f((uint64_t)(h & 0xffffff00u) | (uint32_t)x);
        return;
    }
    f((uint64_t)h);
    for (int i = 0; i < kPayloadBytes[t]; ++i) {
        if (i >= kMinNum[t] && p[i] == 0) break;
        f((uint64_t)(h & 0xffffff00u) | p[i]);
    }
}

bool validate_index(const HostIndex &h, std::string &why)
{
    auto fail = [&](const std::string &m) {
        why = m;
        return false;
    };
    const uint64_t nk = h.nkeys;
    if (h.seg8.size() != (nk + 1) * kSegStride) return fail("seg8 has the wrong size");
    if (h.tile_shift < 8 || h.tile_shift > 16) return fail("value tile out of range");
    if ((uint64_t)h.ntiles != (((uint64_t)h.nvalues + (1ull << h.tile_shift) - 1) >> h.tile_shift))
        return fail("tile count does not match the value space");
    if (h.tile_blk.size() != (size_t)h.ntiles + 1) return fail("tile block list has the wrong size");
    if (h.vkeys.size() < 64) return fail("key stream too short");
    for (int t = 0; t < kTypes; ++t)
        if (h.pay[t].size() != h.hdr[t].size() * kPayloadBytes[t] || h.vhdr[t].size() != h.hdr[t].size() ||
            h.vpay[t].size() != h.pay[t].size())
            return fail("record arrays of a type disagree in size");
    // per-key prefix counts: row 0 zero, every column non-decreasing, the
    // all-type column the sum of the six, totals equal to the arrays
    for (int c = 0; c < kSegStride; ++c)
        if (h.seg8[c] != 0) return fail("seg8 row 0 is not zero");
    for (uint64_t k = 0; k < nk; ++k) {
        const uint32_t *a = &h.seg8[k * kSegStride], *b = a + kSegStride;
        uint64_t sum = 0;
        for (int c = 0; c < kSegStride; ++c)
            if (b[c] < a[c]) return fail("seg8 is not non-decreasing");
        for (int t = 0; t < kTypes; ++t) sum += b[t];
        if (sum != b[6]) return fail("seg8 all-type prefix is not the sum of the types");
    }
    const uint32_t *last = &h.seg8[nk * kSegStride];
    for (int t = 0; t < kTypes; ++t)
        if (last[t] != h.hdr[t].size()) return fail("seg8 totals do not match the record arrays");
    if (last[7] != (uint32_t)h.nnz) return fail("seg8 value total does not match the pair count");
    // every value a record emits lies in the value space; a bitset header is
    // window-aligned; a key's value count is what its records hold (the id
    // bounds of admission rely on it)
    for (int t = 0; t < kTypes; ++t) {
        const uint64_t n = h.hdr[t].size();
        for (uint64_t i = 0; i < n; ++i) {
            bool bad = false;
            auto chk = [&](uint64_t v) { bad = bad || v >= h.nvalues; };
            if (t == AB_BITSET && ((h.hdr[t][i] & 0xffu) || (h.vhdr[t][i] & 0xffu)))
                return fail("bitset header is not window-aligned");
            for_record_values(t, h.hdr[t][i], &h.pay[t][i * kPayloadBytes[t]], chk);
            for_record_values(t, h.vhdr[t][i], &h.vpay[t][i * kPayloadBytes[t]], chk);
            if (bad) return fail("a record holds a value outside the value space");
        }
    }
    {
        uint64_t idx[kTypes] = {0};
        for (uint64_t k = 0; k < nk; ++k) {
            const uint32_t *a = &h.seg8[k * kSegStride], *b = a + kSegStride;
            uint64_t nvals = 0;
            for (int t = 0; t < kTypes; ++t)
                for (; idx[t] < b[t]; ++idx[t])
                    for_record_values(t, h.hdr[t][idx[t]], &h.pay[t][idx[t] * kPayloadBytes[t]],
                                      [&](uint64_t) { ++nvals; });
            if (nvals != (uint64_t)(b[7] - a[7])) return fail("a key's value count disagrees with its records");
        }
    }
    // value-tiled blocks: fields in range, keys and records inside their
    // arrays, every record's window inside the block's tile, every key < nkeys
    if (h.tile_blk[0] != 0 || h.tile_blk.back() != h.vblk.size()) return fail("tile block list ends are wrong");
    for (uint32_t tl = 0; tl < h.ntiles; ++tl)
        if (h.tile_blk[tl + 1] < h.tile_blk[tl]) return fail("tile block list is not non-decreasing");
    for (uint32_t tl = 0; tl < h.ntiles; ++tl) {
        for (uint32_t bi = h.tile_blk[tl]; bi < h.tile_blk[tl + 1]; ++bi) {
            const BlockMeta &m = h.vblk[bi];
            const uint32_t t = m.tw & 15, wl = (m.tw >> 4) & 3, nv = m.tw >> 8;
            if (t >= (uint32_t)kTypes || wl > 2 || (m.tw & 0xc0u) || nv == 0 || nv > kBlockRecs)
                return fail("key-stream block descriptor out of range");
            if ((uint64_t)m.rec_base + nv > h.vhdr[t].size()) return fail("key-stream block records out of range");
            const uint64_t off = (uint64_t)m.off_kb << 10, w = 1ull << wl;
            if (off + (uint64_t)kBlockRecs * w > h.vkeys.size()) return fail("key-stream block bytes out of range");
            for (uint32_t r = 0; r < nv; ++r) {
                const uint8_t *p = &h.vkeys[off + ((uint64_t)(r & 63) * kSlots + (r >> 6)) * w];
                uint32_t d = 0;
                for (uint32_t q = 0; q < w; ++q) d |= (uint32_t)p[q] << (8 * q);
                if ((uint64_t)m.start_key + d >= nk) return fail("key-stream key outside the key space");
                if ((h.vhdr[t][(uint64_t)m.rec_base + r] >> h.tile_shift) != tl)
                    return fail("a block's record lies outside its value tile");
            }
        }
    }
    return true;
}

void host_features(const uint32_t *seg8, uint32_t nkeys, uint32_t nvalues, uint64_t R,
                   const uint32_t *keys, uint32_t n, float *f, uint64_t *m_out)
{
    uint64_t valid = 0, M = 0, mx = 0, bytes = 0, nv = 0, a1 = 0, bs = 0, sq = 0;  // exact integer sums
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k = keys[i];
        if (k >= nkeys) continue;
        const uint32_t *r0 = seg8 + (uint64_t)k * kSegStride, *r1 = r0 + kSegStride;
        uint64_t tot = r1[6] - r0[6];
        ++valid;
        M += tot;
        sq += tot * tot;
        if (tot > mx) mx = tot;
        for (int t = 0; t < kTypes; ++t) bytes += (uint64_t)(r1[t] - r0[t]) * (4 + kPayloadBytes[t]);
        nv += r1[7] - r0[7];
        a1 += r1[0] - r0[0];
        bs += r1[5] - r0[5];
    }
    double dv = valid ? (double)valid : 1.0;
    double mean = (double)M / dv;
    double var = (double)sq / dv - mean * mean;
    if (var < 0) var = 0;
    double dM = M ? (double)M : 1.0;
    double nvd = nvalues ? (double)nvalues : 1.0;
    f[0] = (float)std::log1p((double)n);
    f[1] = (float)std::log1p((double)valid);
    f[2] = (float)std::log1p((double)M);
    f[3] = (float)std::log1p((double)mx);
    f[4] = (float)std::log1p(std::sqrt(var));
    f[5] = (float)std::log1p((double)bytes);
    f[6] = (float)((double)M / (R ? (double)R : 1.0));
    f[7] = (float)std::min(1.0, (double)nv / nvd);
    f[8] = (float)(1.0 - std::exp(-(double)nv / nvd));
    f[9] = (float)((double)valid / (nkeys ? (double)nkeys : 1.0));
    f[10] = (float)((double)a1 / dM);
    f[11] = (float)((double)bs / dM);
    if (m_out) *m_out = M;
}

}  // namespace ab
