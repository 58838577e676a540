/*
 * oracle.c -- CPU restatement of the AdaBool OR-filtering path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never by the product.  See oracle.h for what
 * pins each function.  Reference citations are file:line relative to the
 * reference repository root.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* predicate                                                                 */
/* ------------------------------------------------------------------------ */

int or_csr_build(const uint32_t *keys, const uint32_t *vals, uint64_t nnz,
                 uint32_t nkeys, uint64_t *off, uint32_t *out_vals)
{
    memset(off, 0, sizeof(uint64_t) * ((uint64_t)nkeys + 1));
    for (uint64_t i = 0; i < nnz; ++i) {
        if (keys[i] >= nkeys) return -1;
        off[keys[i] + 1]++;
    }
    for (uint64_t k = 0; k < nkeys; ++k) off[k + 1] += off[k];
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((uint64_t)nkeys + 1));
    if (!cur) return -2;
    memcpy(cur, off, sizeof(uint64_t) * ((uint64_t)nkeys + 1));
    for (uint64_t i = 0; i < nnz; ++i) out_vals[cur[keys[i]]++] = vals[i];
    free(cur);
    return 0;
}

static uint64_t collect_words(const uint32_t *bitmap, uint64_t nwords, uint32_t *ids)
{
    uint64_t c = 0;
    for (uint64_t w = 0; w < nwords; ++w) {
        uint32_t x = bitmap[w];
        while (x) {
            unsigned b = (unsigned)__builtin_ctz(x);
            if (ids) ids[c] = (uint32_t)(w * 32 + b);
            ++c;
            x &= x - 1;
        }
    }
    return c;
}

static uint64_t stride_words(uint32_t nvalues)
{
    uint64_t bytes = ((uint64_t)nvalues + 7) / 8;
    bytes = (bytes + 15) / 16 * 16;
    return bytes / 4;
}

/* test/adsfilter_q_all.cu:196-208: R(q) = {v | (k,v) in filter, k in q}. */
uint64_t or_union_csr(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                      const uint32_t *q, uint32_t n, uint32_t nvalues,
                      uint32_t *bitmap, uint32_t *ids)
{
    uint64_t nw = stride_words(nvalues);
    memset(bitmap, 0, nw * 4);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k = q[i];
        if (k >= nkeys) continue;
        for (uint64_t j = off[k]; j < off[k + 1]; ++j) {
            uint32_t v = vals[j];
            bitmap[v >> 5] |= 1u << (v & 31);
        }
    }
    return collect_words(bitmap, nw, ids);
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* test/adsfilter_q_all.cu:196-208 -- the union as a sorted set of values. */
uint64_t or_union_ids(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                      const uint32_t *q, uint32_t n, uint32_t *ids)
{
    uint64_t m = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k = q[i];
        if (k >= nkeys) continue;
        for (uint64_t j = off[k]; j < off[k + 1]; ++j) ids[m++] = vals[j];
    }
    qsort(ids, m, sizeof(uint32_t), cmp_u32);
    uint64_t u = 0;
    for (uint64_t i = 0; i < m; ++i)
        if (u == 0 || ids[u - 1] != ids[i]) ids[u++] = ids[i];
    return u;
}

/* test/adsfilter_q_all.cu:196-208 -- full COO scan with a query key set. */
uint64_t or_scan_coo(const uint32_t *keys, const uint32_t *vals, uint64_t nnz,
                     const uint32_t *q, uint32_t n, uint32_t nvalues,
                     uint32_t *bitmap)
{
    uint64_t nw = stride_words(nvalues);
    memset(bitmap, 0, nw * 4);
    uint32_t *qs = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    memcpy(qs, q, sizeof(uint32_t) * n);
    qsort(qs, n, sizeof(uint32_t), cmp_u32);
    for (uint64_t i = 0; i < nnz; ++i) {
        uint32_t k = keys[i];
        /* set membership by binary search (std::set::count) */
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (qs[mid] < k) lo = mid + 1; else hi = mid;
        }
        if (lo < n && qs[lo] == k) {
            uint32_t v = vals[i];
            bitmap[v >> 5] |= 1u << (v & 31);
        }
    }
    free(qs);
    return collect_words(bitmap, nw, NULL);
}

/* CollectIdsFromBitmap, test/adsfilter_batch.cu:82-97 (byte-wise, LSB first). */
uint64_t or_collect_ids(const uint32_t *bitmap, uint64_t nbits, uint32_t limit,
                        uint32_t *ids)
{
    const uint8_t *bytes = (const uint8_t *)bitmap;
    uint64_t total = nbits < limit ? nbits : limit;
    uint64_t c = 0;
    for (uint64_t b = 0; b < (total + 7) / 8; ++b) {
        unsigned x = bytes[b];
        while (x) {
            unsigned lsb = (unsigned)__builtin_ctz(x);
            uint64_t id = b * 8 + lsb;
            if (id < limit) {
                if (ids) ids[c] = (uint32_t)id;
                ++c;
            }
            x &= x - 1;
        }
    }
    return c;
}

/* ------------------------------------------------------------------------ */
/* reference index format                                                    */
/* ------------------------------------------------------------------------ */

static const int kPayloadBytes[OR_NTYPES] = {0, 4, 8, 16, 32, 32};
int or_payload_bytes(int t) { return (t >= 0 && t < OR_NTYPES) ? kPayloadBytes[t] : -1; }

typedef struct {
    uint8_t *p;
    uint64_t n, cap;
} vec8;
typedef struct {
    uint32_t *p;
    uint64_t n, cap;
} vec32;

static void v8_push(vec8 *v, uint8_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->p = (uint8_t *)realloc(v->p, v->cap);
    }
    v->p[v->n++] = x;
}
static void v32_push(vec32 *v, uint32_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->p = (uint32_t *)realloc(v->p, v->cap * 4);
    }
    v->p[v->n++] = x;
}

struct or_ref_index {
    vec32 headers[OR_NTYPES];
    vec8 payload[OR_NTYPES];
    vec32 keys[OR_NTYPES];
    vec32 segs[OR_NTYPES];
};

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Type of a (key, window) bucket of n low bytes:
 * src/AdaBool/inverted_list.h:425-458 (<=1, <=5, <=9, <=17, <=33, else). */
static int bucket_type(uint64_t n)
{
    if (n <= 1) return OR_ARRAY1;
    if (n <= 5) return OR_ARRAY4;
    if (n <= 9) return OR_ARRAY8;
    if (n <= 17) return OR_ARRAY16;
    if (n <= 33) return OR_ARRAY32;
    return OR_BITSET;
}

/* ivf_list::BuildFrom + GroupKeys, src/AdaBool/inverted_list.h:348-470.
 * Buckets are visited in (key, window) order like the nested std::map; the
 * low bytes of a bucket are sorted (AddArray, :407). */
or_ref_index *or_ref_index_build(const uint32_t *keys, const uint32_t *vals,
                                 uint64_t nnz, int dedupe)
{
    or_ref_index *ix = (or_ref_index *)calloc(1, sizeof(or_ref_index));
    uint64_t *kv = (uint64_t *)malloc(sizeof(uint64_t) * (nnz ? nnz : 1));
    for (uint64_t i = 0; i < nnz; ++i) kv[i] = ((uint64_t)keys[i] << 32) | vals[i];
    qsort(kv, nnz, sizeof(uint64_t), cmp_u64);
    if (dedupe && nnz) {
        uint64_t w = 1;
        for (uint64_t i = 1; i < nnz; ++i)
            if (kv[i] != kv[w - 1]) kv[w++] = kv[i];
        nnz = w;
    }
    for (int t = 0; t < OR_NTYPES; ++t) v32_push(&ix->segs[t], 0);

    uint64_t i = 0;
    while (i < nnz) {
        uint32_t key = (uint32_t)(kv[i] >> 32);
        uint64_t cnt[OR_NTYPES] = {0};
        while (i < nnz && (uint32_t)(kv[i] >> 32) == key) {
            uint32_t win = (uint32_t)kv[i] & 0xffffff00u;
            uint64_t j = i;
            while (j < nnz && (uint32_t)(kv[j] >> 32) == key &&
                   ((uint32_t)kv[j] & 0xffffff00u) == win)
                ++j;
            uint64_t n = j - i; /* sorted low bytes are kv[i..j) & 0xff */
            int t = bucket_type(n);
            cnt[t]++;
            v32_push(&ix->keys[t], key);
            if (t == OR_BITSET) {
                uint8_t bits[32] = {0};
                for (uint64_t b = i; b < j; ++b) {
                    uint32_t lo = (uint32_t)kv[b] & 0xff;
                    bits[lo / 8] |= (uint8_t)(1u << (lo % 8));
                }
                v32_push(&ix->headers[t], win);
                for (int b = 0; b < 32; ++b) v8_push(&ix->payload[t], bits[b]);
            } else {
                v32_push(&ix->headers[t], win | ((uint32_t)kv[i] & 0xff));
                int plen = kPayloadBytes[t];
                for (int b = 0; b < plen; ++b) {
                    uint8_t x = (uint64_t)(b + 1) < n ? (uint8_t)(kv[i + b + 1] & 0xff) : 0;
                    v8_push(&ix->payload[t], x);
                }
            }
            i = j;
        }
        /* key_segments_ prefix, src/AdaBool/inverted_list.h:461-468 */
        for (int t = 0; t < OR_NTYPES; ++t) {
            vec32 *s = &ix->segs[t];
            while (s->n <= key) v32_push(s, s->p[s->n - 1]);
            v32_push(s, s->p[s->n - 1] + (uint32_t)cnt[t]);
        }
    }
    free(kv);
    return ix;
}

void or_ref_index_free(or_ref_index *ix)
{
    if (!ix) return;
    for (int t = 0; t < OR_NTYPES; ++t) {
        free(ix->headers[t].p);
        free(ix->payload[t].p);
        free(ix->keys[t].p);
        free(ix->segs[t].p);
    }
    free(ix);
}

uint64_t or_ref_index_nrec(const or_ref_index *ix, int t) { return ix->headers[t].n; }
const uint32_t *or_ref_index_headers(const or_ref_index *ix, int t) { return ix->headers[t].p; }
const uint8_t *or_ref_index_payload(const or_ref_index *ix, int t) { return ix->payload[t].p; }
const uint32_t *or_ref_index_keys(const or_ref_index *ix, int t) { return ix->keys[t].p; }
const uint32_t *or_ref_index_segs(const or_ref_index *ix, int t, uint64_t *seglen)
{
    if (seglen) *seglen = ix->segs[t].n;
    return ix->segs[t].p;
}

/* ComputeKeyBlockMeta, src/AdaBool/inverted_list.h:90-148.  The span is taken
 * as an unsigned 32-bit difference (the reference stores it in an int, which
 * only differs for spans >= 2^31). */
uint64_t or_key_block_meta(uint32_t step, uint64_t len, const uint32_t *keys,
                           uint32_t *starts, uint32_t *pos, uint8_t *buf)
{
    if (len == 0 || step == 0) return 0;
    uint64_t nblocks = (len + step - 1) / step;
    uint64_t bytes = 0;
    uint32_t p = 0;
    if (pos) pos[0] = 0;
    for (uint64_t b = 0; b < nblocks; ++b) {
        uint32_t beg = keys[b * step];
        uint32_t nxt = (b + 1 < nblocks) ? keys[(b + 1) * step] : keys[len - 1];
        uint32_t span = nxt - beg;
        uint32_t w = span < 256u ? 1 : (span < 65536u ? 2 : 4);
        uint64_t num = len - b * step < step ? len - b * step : step;
        if (starts) starts[b] = beg;
        p += (uint32_t)(num * w);
        if (pos) pos[b + 1] = p;
        if (buf) {
            for (uint64_t i = 0; i < num; ++i) {
                uint32_t d = keys[b * step + i] - beg;
                for (uint32_t j = 0; j < w; ++j) buf[bytes + i * w + j] = (uint8_t)(d >> (8 * j));
            }
            memset(buf + bytes + num * w, 0, (step - num) * w);
        }
        bytes += (uint64_t)step * w;
    }
    if (starts) starts[nblocks] = keys[len - 1];
    return bytes;
}

/* Decode of src/AdaBool/inverted_list.h:686-747: ARRAY records stop at the
 * first zero payload byte; BITSET records list their set bits. */
uint64_t or_ref_index_decode(const or_ref_index *ix, uint32_t *out_keys, uint32_t *out_vals)
{
    uint64_t c = 0;
    for (int t = 0; t < OR_NTYPES; ++t) {
        int plen = kPayloadBytes[t];
        for (uint64_t r = 0; r < ix->headers[t].n; ++r) {
            uint32_t h = ix->headers[t].p[r];
            uint32_t key = ix->keys[t].p[r];
            uint32_t hi = h & 0xffffff00u;
            const uint8_t *pl = ix->payload[t].p + r * plen;
            if (t == OR_BITSET) {
                for (int b = 0; b < 256; ++b)
                    if (pl[b / 8] & (1u << (b % 8))) {
                        if (out_keys) { out_keys[c] = key; out_vals[c] = hi | (uint32_t)b; }
                        ++c;
                    }
            } else {
                if (out_keys) { out_keys[c] = key; out_vals[c] = h; }
                ++c;
                for (int j = 0; j < plen; ++j) {
                    if (pl[j] == 0) break;
                    if (out_keys) { out_keys[c] = key; out_vals[c] = hi | pl[j]; }
                    ++c;
                }
            }
        }
    }
    return c;
}

/* kMinNum of ArraySizeTrait, src/AdaBool/inverted_list.cu.h:21-52. */
static const int kMinNum[OR_NTYPES] = {0, 1, 5, 9, 17, 0};

void or_emit(int t, uint32_t header, const uint8_t *payload, uint32_t *bitmap)
{
    if (t == OR_BITSET) {
        /* src/AdaBool/key_transformer.cu.h:100-111: word (header>>5)+j |= bits[j] */
        for (uint32_t j = 0; j < 8; ++j) {
            uint32_t w = (uint32_t)payload[4 * j] | ((uint32_t)payload[4 * j + 1] << 8) |
                         ((uint32_t)payload[4 * j + 2] << 16) | ((uint32_t)payload[4 * j + 3] << 24);
            bitmap[((header | (j * 32)) >> 5)] |= w;
        }
        return;
    }
    bitmap[header >> 5] |= 1u << (header & 31);
    if (t == OR_ARRAY1) return;
    uint32_t h = header & 0xffffff00u;
    int size = kPayloadBytes[t];
    for (int i = 0; i < size; ++i) {
        if (i >= kMinNum[t] && payload[i] == 0) return;
        uint32_t v = h | payload[i];
        bitmap[v >> 5] |= 1u << (v & 31);
    }
}

/* ListUnion_dense, test/adsfilter_q_all.cu:26-61; per-block decode
 * ForBPKeyIdxs, src/AdaBool/key_transformer.cu.h:11-44; bit test :75-82. */
void or_dense_probe(const or_ref_index *ix, uint32_t step, const uint32_t *qbitmap,
                    uint64_t nq_bits, uint32_t *bitmap)
{
    for (int t = 0; t < OR_NTYPES; ++t) {
        uint64_t len = ix->headers[t].n;
        if (!len) continue;
        uint64_t nblocks = (len + step - 1) / step;
        uint32_t *starts = (uint32_t *)malloc(4 * (nblocks + 1));
        uint32_t *pos = (uint32_t *)malloc(4 * (nblocks + 1));
        uint64_t nbytes = or_key_block_meta(step, len, ix->keys[t].p, NULL, NULL, NULL);
        uint8_t *buf = (uint8_t *)malloc(nbytes ? nbytes : 1);
        or_key_block_meta(step, len, ix->keys[t].p, starts, pos, buf);
        int plen = kPayloadBytes[t];
        for (uint64_t b = 0; b < nblocks; ++b) {
            uint32_t kb = starts[b];
            uint32_t span = starts[b + 1] - kb;
            uint32_t w = span < 256u ? 1 : (span < 65536u ? 2 : 4);
            const uint8_t *kbuf = buf + pos[b];
            for (uint32_t i = 0; i < step; ++i) {
                uint64_t idx = b * step + i;
                if (idx >= len) continue;
                uint32_t d = 0;
                for (uint32_t j = 0; j < w; ++j) d |= (uint32_t)kbuf[i * w + j] << (8 * j);
                uint32_t key = kb + d;
                if (key >= nq_bits) continue;
                if (qbitmap[key >> 5] & (1u << (key & 31)))
                    or_emit(t, ix->headers[t].p[idx], ix->payload[t].p + idx * plen, bitmap);
            }
        }
        free(starts);
        free(pos);
        free(buf);
    }
}

/* Prepare(keys), src/AdaBool/inverted_list.h:481-527, chunked by
 * LoadBalanceTransformer::Prepare, src/AdaBool/load_balance_transformer.h:50-69;
 * each unit emitted like ListUnionKernel*, load_balance_transformer.cu.h:16-91. */
uint64_t or_sparse_union(const or_ref_index *ix, const uint32_t *q, uint32_t n,
                         uint32_t step, uint32_t *bitmap)
{
    uint64_t units = 0;
    for (int t = 0; t < OR_NTYPES; ++t) {
        const uint32_t *segs = ix->segs[t].p;
        uint64_t seglen = ix->segs[t].n;
        int plen = kPayloadBytes[t];
        for (uint32_t s = 0; s < n; ++s) {
            uint32_t key = q[s];
            uint32_t beg = 0, len = 0;
            if ((uint64_t)key < seglen - 1) {
                beg = segs[key];
                len = segs[key + 1] - beg;
            }
            for (uint32_t off = 0; off < len; off += step) {
                uint32_t ul = len - off < step ? len - off : step;
                ++units;
                for (uint32_t i = 0; i < ul; ++i) {
                    uint64_t idx = (uint64_t)beg + off + i;
                    or_emit(t, ix->headers[t].p[idx], ix->payload[t].p + idx * plen, bitmap);
                }
            }
        }
    }
    return units;
}

/* ------------------------------------------------------------------------ */
/* multi-threaded CPU baseline                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    const uint64_t *off;
    const uint32_t *vals;
    uint32_t nkeys;
    const uint32_t *qkeys;
    const uint64_t *qoff;
    uint32_t nq;
    uint32_t nvalues;
    uint64_t *counts;
    double *lat_ms;
    volatile uint32_t next;
    uint64_t checksum;
    pthread_mutex_t mu;
} batch_job;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void *batch_worker(void *arg)
{
    batch_job *J = (batch_job *)arg;
    uint64_t nw = stride_words(J->nvalues);
    uint32_t *bm = (uint32_t *)malloc(nw * 4);
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * ((uint64_t)J->nvalues + 1));
    uint64_t local = 0;
    for (;;) {
        uint32_t qi = __sync_fetch_and_add(&J->next, 1);
        if (qi >= J->nq) break;
        double t0 = now_s();
        uint64_t c = or_union_csr(J->off, J->vals, J->nkeys, J->qkeys + J->qoff[qi],
                                  (uint32_t)(J->qoff[qi + 1] - J->qoff[qi]), J->nvalues, bm, ids);
        double t1 = now_s();
        J->counts[qi] = c;
        if (J->lat_ms) J->lat_ms[qi] = (t1 - t0) * 1e3;
        uint64_t h = c * 0x9E3779B97F4A7C15ull;
        for (uint64_t i = 0; i < c; i += 997) h ^= (uint64_t)ids[i] * (i + 1);
        local += h;
    }
    pthread_mutex_lock(&J->mu);
    J->checksum += local;
    pthread_mutex_unlock(&J->mu);
    free(bm);
    free(ids);
    return NULL;
}

double or_union_batch(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                      const uint32_t *qkeys, const uint64_t *qoff, uint32_t nq,
                      uint32_t nvalues, int nthreads, uint64_t *counts,
                      double *lat_ms, uint64_t *checksum)
{
    batch_job J;
    memset(&J, 0, sizeof(J));
    J.off = off; J.vals = vals; J.nkeys = nkeys; J.qkeys = qkeys; J.qoff = qoff;
    J.nq = nq; J.nvalues = nvalues; J.counts = counts; J.lat_ms = lat_ms;
    pthread_mutex_init(&J.mu, NULL);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    double t0 = now_s();
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, batch_worker, &J);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    double t1 = now_s();
    free(th);
    pthread_mutex_destroy(&J.mu);
    if (checksum) *checksum = J.checksum;
    return t1 - t0;
}

/* ------------------------------------------------------------------------ */
/* per-query digests (checkers for large batches)                            */
/* ------------------------------------------------------------------------ */

static inline uint64_t dg_wt(uint64_t w) { return ((w + 1) * 0x9E3779B97F4A7C15ull) | 1ull; }
static inline uint64_t dg_mix(uint32_t x)
{
    return ((uint64_t)x * 0x9E3779B1ull) ^ ((uint64_t)(x ^ 0x5BD1E995u) * 0x27D4EB2Full);
}

static uint64_t dg_bitmap(const uint32_t *bm, uint64_t nwords)
{
    uint64_t h = 0;
    for (uint64_t w = 0; w < nwords; ++w) h += (uint64_t)bm[w] * dg_wt(w);
    return h;
}

static uint64_t dg_ids(const uint32_t *ids, uint64_t n, uint8_t *asc)
{
    uint64_t h = 0;
    uint8_t ok = 1;
    for (uint64_t i = 0; i < n; ++i) {
        h += dg_mix(ids[i]);
        if (i && ids[i] <= ids[i - 1]) ok = 0;
    }
    if (asc) *asc = ok;
    return h;
}

typedef struct {
    int kind; /* 0 union, 1 bitmaps, 2 ids */
    const uint64_t *off;
    const uint32_t *vals;
    uint32_t nkeys, nvalues;
    const uint32_t *qkeys;
    const uint64_t *qoff;
    const uint32_t *data;
    uint64_t stride, nwords;
    const uint64_t *doff;
    const uint32_t *dcnt;
    uint64_t nq;
    uint64_t *counts, *d0, *d1;
    uint8_t *asc;
    volatile uint64_t next;
} digest_job;

static void *digest_worker(void *arg)
{
    digest_job *J = (digest_job *)arg;
    uint32_t *bm = NULL, *ids = NULL;
    uint64_t nw = 0;
    if (J->kind == 0) {
        nw = stride_words(J->nvalues);
        bm = (uint32_t *)malloc(nw * 4 + 4);
        ids = (uint32_t *)malloc(sizeof(uint32_t) * ((uint64_t)J->nvalues + 1));
    }
    for (;;) {
        uint64_t q = __sync_fetch_and_add(&J->next, 1);
        if (q >= J->nq) break;
        if (J->kind == 0) {
            uint64_t c = or_union_csr(J->off, J->vals, J->nkeys, J->qkeys + J->qoff[q],
                                      (uint32_t)(J->qoff[q + 1] - J->qoff[q]), J->nvalues, bm, ids);
            J->counts[q] = c;
            J->d0[q] = dg_bitmap(bm, nw);
            J->d1[q] = dg_ids(ids, c, NULL);
        } else if (J->kind == 1) {
            J->d0[q] = dg_bitmap(J->data + q * J->stride, J->nwords);
        } else {
            J->d0[q] = dg_ids(J->data + J->doff[q], J->dcnt[q], J->asc ? &J->asc[q] : NULL);
        }
    }
    free(bm);
    free(ids);
    return NULL;
}

static void run_digest(digest_job *J, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, digest_worker, J);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
}

void or_union_batch_digest(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                           const uint32_t *qkeys, const uint64_t *qoff, uint32_t nq,
                           uint32_t nvalues, int nthreads, uint64_t *counts,
                           uint64_t *bm_digest, uint64_t *id_digest)
{
    digest_job J;
    memset(&J, 0, sizeof(J));
    J.kind = 0; J.off = off; J.vals = vals; J.nkeys = nkeys; J.nvalues = nvalues;
    J.qkeys = qkeys; J.qoff = qoff; J.nq = nq; J.counts = counts; J.d0 = bm_digest; J.d1 = id_digest;
    run_digest(&J, nthreads);
}

void or_bitmap_digest(const uint32_t *bm, uint64_t stride_words, uint64_t nwords,
                      uint64_t nq, int nthreads, uint64_t *out)
{
    digest_job J;
    memset(&J, 0, sizeof(J));
    J.kind = 1; J.data = bm; J.stride = stride_words; J.nwords = nwords; J.nq = nq; J.d0 = out;
    run_digest(&J, nthreads);
}

void or_ids_digest(const uint32_t *ids, const uint64_t *off, const uint32_t *counts,
                   uint64_t nq, int nthreads, uint64_t *out, uint8_t *ascending)
{
    digest_job J;
    memset(&J, 0, sizeof(J));
    J.kind = 2; J.data = ids; J.doff = off; J.dcnt = counts; J.nq = nq; J.d0 = out; J.asc = ascending;
    run_digest(&J, nthreads);
}

/* ------------------------------------------------------------------------ */
/* statistics                                                                */
/* ------------------------------------------------------------------------ */

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* CalculateStats, test/adsfilter_q_all.cu:258-307. */
void or_stats(double *d, uint64_t n, double *o)
{
    memset(o, 0, sizeof(double) * 9);
    if (!n) return;
    qsort(d, n, sizeof(double), cmp_double);
    double s = 0;
    for (uint64_t i = 0; i < n; ++i) s += d[i];
    double mean = s / n;
    double median = (n % 2 == 0) ? (d[n / 2 - 1] + d[n / 2]) / 2.0 : d[n / 2];
    double sq = 0;
    for (uint64_t i = 0; i < n; ++i) sq += (d[i] - mean) * (d[i] - mean);
    o[0] = mean;
    o[1] = median;
    o[2] = sqrt(sq / n);
    o[3] = d[0];
    o[4] = d[n - 1];
    o[5] = median;
    const double ps[3] = {0.90, 0.95, 0.99};
    for (int k = 0; k < 3; ++k) {
        double p = ps[k] * (double)(n - 1);
        uint64_t lo = (uint64_t)p, up = lo + 1;
        o[6 + k] = (up >= n) ? d[n - 1] : d[lo] * (1 - (p - lo)) + d[up] * (p - lo);
    }
}

/* ------------------------------------------------------------------------ */
/* text formats                                                              */
/* ------------------------------------------------------------------------ */

static int is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

/* Parses one unsigned 32-bit number at *s after skipping blanks, like
 * `istream >> uint32_t`; a sign or overflow is a failure. */
static int parse_u32(const char **s, uint32_t *out)
{
    const char *p = *s;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p < '0' || *p > '9') return 0;
    uint64_t v = 0;
    while (*p >= '0' && *p <= '9') {
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xffffffffull) return 0;
        ++p;
    }
    *out = (uint32_t)v;
    *s = p;
    return 1;
}

/* Trims and classifies a line (test/adsfilter_q_all.cu:138-146). */
static const char *content_of(char *line)
{
    size_t L = strlen(line);
    while (L && is_ws(line[L - 1])) line[--L] = 0;
    char *p = line;
    while (*p && is_ws(*p)) ++p;
    if (!*p || *p == '#' || *p == '%') return NULL;
    return p;
}

int64_t or_read_coo(const char *path, uint32_t *keys, uint32_t *vals,
                    uint32_t *max_row, uint32_t *max_col)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char *line = NULL;
    size_t cap = 0;
    int64_t n = 0;
    uint32_t mr = 0, mc = 0;
    while (getline(&line, &cap, f) >= 0) {
        const char *p = content_of(line);
        if (!p) continue;
        uint32_t r, c;
        if (!parse_u32(&p, &r)) continue;
        if (!parse_u32(&p, &c)) continue;
        if (keys) { keys[n] = r; vals[n] = c; }
        if (r > mr) mr = r;
        if (c > mc) mc = c;
        ++n;
    }
    free(line);
    fclose(f);
    if (max_row) *max_row = mr;
    if (max_col) *max_col = mc;
    return n;
}

int64_t or_read_query(const char *path, uint32_t *keys)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char *line = NULL;
    size_t cap = 0;
    int64_t n = 0;
    while (getline(&line, &cap, f) >= 0) {
        const char *p = content_of(line);
        if (!p) continue;
        uint32_t r;
        if (!parse_u32(&p, &r)) continue;
        if (keys) keys[n] = r;
        ++n;
    }
    free(line);
    fclose(f);
    return n;
}

/* ------------------------------------------------------------------------ */
/* MLP cost model reference (float64)                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    const double *W1, *b1, *W2, *b2, *W3, *b3;
} mlp_view;

static mlp_view mlp_split(const double *p, int F, int H1, int H2, int O)
{
    mlp_view m;
    m.W1 = p; p += (size_t)H1 * F;
    m.b1 = p; p += H1;
    m.W2 = p; p += (size_t)H2 * H1;
    m.b2 = p; p += H2;
    m.W3 = p; p += (size_t)O * H2;
    m.b3 = p;
    return m;
}

static void mlp_fwd1(mlp_view m, int F, int H1, int H2, int O, const double *x,
                     double *z1, double *h1, double *z2, double *h2, double *y)
{
    for (int j = 0; j < H1; ++j) {
        double s = m.b1[j];
        for (int i = 0; i < F; ++i) s += m.W1[(size_t)j * F + i] * x[i];
        z1[j] = s;
        h1[j] = s > 0 ? s : 0;
    }
    for (int j = 0; j < H2; ++j) {
        double s = m.b2[j];
        for (int i = 0; i < H1; ++i) s += m.W2[(size_t)j * H1 + i] * h1[i];
        z2[j] = s;
        h2[j] = s > 0 ? s : 0;
    }
    for (int j = 0; j < O; ++j) {
        double s = m.b3[j];
        for (int i = 0; i < H2; ++i) s += m.W3[(size_t)j * H2 + i] * h2[i];
        y[j] = s;
    }
}

void or_mlp_forward(const double *params, int F, int H1, int H2, int O,
                    const double *x, int n, double *y)
{
    mlp_view m = mlp_split(params, F, H1, H2, O);
    double *z1 = (double *)malloc(sizeof(double) * (H1 + H1 + H2 + H2));
    double *h1 = z1 + H1, *z2 = h1 + H1, *h2 = z2 + H2;
    for (int r = 0; r < n; ++r) mlp_fwd1(m, F, H1, H2, O, x + (size_t)r * F, z1, h1, z2, h2, y + (size_t)r * O);
    free(z1);
}

double or_mlp_loss_grad(const double *params, int F, int H1, int H2, int O,
                        const double *x, const double *y, int n, double *grad)
{
    mlp_view m = mlp_split(params, F, H1, H2, O);
    size_t np = (size_t)H1 * F + H1 + (size_t)H2 * H1 + H2 + (size_t)O * H2 + O;
    memset(grad, 0, sizeof(double) * np);
    double *gW1 = grad, *gb1 = gW1 + (size_t)H1 * F, *gW2 = gb1 + H1, *gb2 = gW2 + (size_t)H2 * H1;
    double *gW3 = gb2 + H2, *gb3 = gW3 + (size_t)O * H2;
    double *buf = (double *)malloc(sizeof(double) * (3 * H1 + 3 * H2 + 2 * O));
    double *z1 = buf, *h1 = z1 + H1, *d1 = h1 + H1, *z2 = d1 + H1, *h2 = z2 + H2, *d2 = h2 + H2;
    double *yh = d2 + H2, *dy = yh + O;
    double loss = 0, scale = 1.0 / ((double)n * O);
    for (int r = 0; r < n; ++r) {
        const double *xr = x + (size_t)r * F;
        mlp_fwd1(m, F, H1, H2, O, xr, z1, h1, z2, h2, yh);
        for (int j = 0; j < O; ++j) {
            double e = yh[j] - y[(size_t)r * O + j];
            loss += 0.5 * e * e * scale;
            dy[j] = e * scale;
            gb3[j] += dy[j];
            for (int i = 0; i < H2; ++i) gW3[(size_t)j * H2 + i] += dy[j] * h2[i];
        }
        for (int i = 0; i < H2; ++i) {
            double s = 0;
            for (int j = 0; j < O; ++j) s += m.W3[(size_t)j * H2 + i] * dy[j];
            d2[i] = z2[i] > 0 ? s : 0;
            gb2[i] += d2[i];
            for (int k = 0; k < H1; ++k) gW2[(size_t)i * H1 + k] += d2[i] * h1[k];
        }
        for (int k = 0; k < H1; ++k) {
            double s = 0;
            for (int i = 0; i < H2; ++i) s += m.W2[(size_t)i * H1 + k] * d2[i];
            d1[k] = z1[k] > 0 ? s : 0;
            gb1[k] += d1[k];
            for (int f = 0; f < F; ++f) gW1[(size_t)k * F + f] += d1[k] * xr[f];
        }
    }
    free(buf);
    return loss;
}

void or_adam_step(double *p, double *m, double *v, const double *g, int np, double lr,
                  double b1, double b2, double eps, int step)
{
    double c1 = 1.0 - pow(b1, step), c2 = 1.0 - pow(b2, step);
    for (int i = 0; i < np; ++i) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
        double mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (sqrt(vh) + eps);
    }
}
