/*
 * oracle.h -- CPU restatement of the AdaBool Boolean OR-filtering path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (adabool_amd/) never links, loads or calls anything here.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Plain C99, no dependencies beyond libc and
 * pthreads.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_union_csr        -- brute-force COO scan (or_scan_coo), both restated
 *                          from test/adsfilter_q_all.cu:196-208; hand-checked
 *                          fixtures in tests/golden/.
 *   or_ref_index_build  -- decode(index) == CSR; bucket-type thresholds and
 *                          payload bytes checked against hand-built records.
 *   or_key_block_meta   -- known-answer example of
 *                          src/AdaBool/inverted_list.h:73-81.
 *   or_dense_probe      -- equals or_union_csr on random inputs (different
 *                          algorithm: key-stream scan + emit vs. CSR walk).
 *   or_sparse_union     -- equals or_union_csr (extent lookup + LBS chunking
 *                          + emit vs. CSR walk).
 *   or_union_ids        -- equals or_union_csr's ids on random inputs
 *                          (gather + sort + dedupe vs. bitmap walk) and a
 *                          pure-Python set union on tiny ones, including
 *                          values at 2^32 - 2.
 *   or_*_digest         -- or_union_batch_digest equals or_bitmap_digest /
 *                          or_ids_digest of or_union_csr's own results, and
 *                          a pure-Python evaluation of the two sums on
 *                          small inputs; one flipped bit or id changes them.
 *   or_stats            -- closed forms (median of even/odd n, linear
 *                          interpolation) checked against numpy.percentile.
 *   or_mlp_*            -- finite differences and a closed-form 1-layer case.
 */
#ifndef ADABOOL_ORACLE_H
#define ADABOOL_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_ARRAY1 = 0, OR_ARRAY4, OR_ARRAY8, OR_ARRAY16, OR_ARRAY32, OR_BITSET, OR_NTYPES };

/* ---- predicate ---------------------------------------------------------- */
/* CSR by key: off[nkeys+1], vals[off[nkeys]].  Built by counting sort. */
int or_csr_build(const uint32_t *keys, const uint32_t *vals, uint64_t nnz,
                 uint32_t nkeys, uint64_t *off, uint32_t *out_vals);

/* Union of the lists of q[0..n) into bitmap (nwords u32, zeroed here); ids in
 * ascending order into ids (may be NULL); returns |R(q)|.  Keys >= nkeys add
 * nothing.  test/adsfilter_q_all.cu:196-208 (predicate), 108-119 (ids). */
uint64_t or_union_csr(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                      const uint32_t *q, uint32_t n, uint32_t nvalues,
                      uint32_t *bitmap, uint32_t *ids);

/* The same set without a bitmap: the lists of q's in-range keys gathered
 * into ids (room for the sum of their lengths), sorted and deduplicated in
 * place; returns |R(q)|.  test/adsfilter_q_all.cu:196-208 (predicate),
 * 108-119 (ascending ids). */
uint64_t or_union_ids(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                      const uint32_t *q, uint32_t n, uint32_t *ids);

/* Brute-force scan over all pairs (the reference's own ground truth,
 * test/adsfilter_q_all.cu:196-208). */
uint64_t or_scan_coo(const uint32_t *keys, const uint32_t *vals, uint64_t nnz,
                     const uint32_t *q, uint32_t n, uint32_t nvalues,
                     uint32_t *bitmap);

/* Bitmap -> ascending ids, restating CollectIdsFromBitmap
 * (test/adsfilter_batch.cu:82-97): bytes LSB-first, ids >= limit dropped. */
uint64_t or_collect_ids(const uint32_t *bitmap, uint64_t nbits, uint32_t limit,
                        uint32_t *ids);

/* ---- reference index format --------------------------------------------- */
typedef struct or_ref_index or_ref_index;
/* Restates ivf_list::BuildFrom + GroupKeys (src/AdaBool/inverted_list.h:
 * 348-470).  dedupe=0 keeps duplicate pairs exactly as the reference does
 * (its buckets count duplicates); dedupe=1 removes them first. */
or_ref_index *or_ref_index_build(const uint32_t *keys, const uint32_t *vals,
                                 uint64_t nnz, int dedupe);
void or_ref_index_free(or_ref_index *ix);
/* Per type: number of records, header/payload/key arrays, key_segments
 * (length seglen = max_key+2). */
uint64_t or_ref_index_nrec(const or_ref_index *ix, int t);
const uint32_t *or_ref_index_headers(const or_ref_index *ix, int t);
const uint8_t *or_ref_index_payload(const or_ref_index *ix, int t);
const uint32_t *or_ref_index_keys(const or_ref_index *ix, int t);
const uint32_t *or_ref_index_segs(const or_ref_index *ix, int t, uint64_t *seglen);
int or_payload_bytes(int t);

/* ComputeKeyBlockMeta (src/AdaBool/inverted_list.h:90-148).  Sizes first:
 * nblocks = ceil(len/step); starts/pos have nblocks+1 entries; returns the
 * padded buffer byte count.  Fills the arrays when they are non-NULL. */
uint64_t or_key_block_meta(uint32_t step, uint64_t len, const uint32_t *keys,
                           uint32_t *starts, uint32_t *pos, uint8_t *buf);

/* Decode every record of the index into (key, value) pairs, restating the
 * decode of src/AdaBool/inverted_list.h:686-747; returns the pair count. */
uint64_t or_ref_index_decode(const or_ref_index *ix, uint32_t *out_keys,
                             uint32_t *out_vals);

/* Emit one record into a bitmap: ProcessArrayFunctor
 * (src/AdaBool/inverted_list.cu.h:61-94), ARRAY1 inline
 * (src/AdaBool/key_transformer.cu.h:297-298), bitset
 * (src/AdaBool/key_transformer.cu.h:100-111). */
void or_emit(int t, uint32_t header, const uint8_t *payload, uint32_t *bitmap);

/* Dense bitmap-probe over the key stream: ListUnion_dense
 * (test/adsfilter_q_all.cu:26-61) with KeyBP kernels
 * (src/AdaBool/key_transformer.cu.h:11-331).  qbitmap holds ceil(nq_bits/32)
 * words; bitmap is OR-ed (caller zeroes). */
void or_dense_probe(const or_ref_index *ix, uint32_t step,
                    const uint32_t *qbitmap, uint64_t nq_bits, uint32_t *bitmap);

/* Sparse list union: Prepare(keys) (src/AdaBool/inverted_list.h:481-527),
 * LoadBalanceTransformer::Prepare (src/AdaBool/load_balance_transformer.h:
 * 50-69) and ListUnionKernel* (src/AdaBool/load_balance_transformer.cu.h:
 * 16-91).  Returns the number of work units; bitmap is OR-ed. */
uint64_t or_sparse_union(const or_ref_index *ix, const uint32_t *q, uint32_t n,
                         uint32_t step, uint32_t *bitmap);

/* ---- batch / baseline ----------------------------------------------------- */
/* Union for nq queries (keys concatenated, qoff[nq+1]) with nthreads worker
 * threads, one query per task; writes counts[nq] and per-query latency in
 * ms (lat_ms may be NULL).  Returns wall seconds. */
double or_union_batch(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                      const uint32_t *qkeys, const uint64_t *qoff, uint32_t nq,
                      uint32_t nvalues, int nthreads, uint64_t *counts,
                      double *lat_ms, uint64_t *checksum);

/* Per-query digests, so that large batches can be compared query by query
 * without holding every result twice.  A bitmap's digest is
 *   sum over words w of bm[w] * wt(w)  (mod 2^64),  wt(w) = (w+1)*G | 1,
 * G = 0x9E3779B97F4A7C15; an id list's digest is
 *   sum over ids x of mix(x)  (mod 2^64),
 *   mix(x) = x*0x9E3779B1 ^ (x ^ 0x5BD1E995)*0x27D4EB2F (64-bit products),
 * and it is marked not ascending when some id is <= the one before it.
 * (Checkers only: two results with equal count and digests are taken to be
 * equal.)  or_union_batch_digest: the oracle's union of every query
 * (or_union_csr) reduced to (count, bitmap digest, id digest), nthreads
 * workers. */
void or_union_batch_digest(const uint64_t *off, const uint32_t *vals, uint32_t nkeys,
                           const uint32_t *qkeys, const uint64_t *qoff, uint32_t nq,
                           uint32_t nvalues, int nthreads, uint64_t *counts,
                           uint64_t *bm_digest, uint64_t *id_digest);
/* Digests of nq bitmaps of nwords words at bm + q * stride_words. */
void or_bitmap_digest(const uint32_t *bm, uint64_t stride_words, uint64_t nwords,
                      uint64_t nq, int nthreads, uint64_t *out);
/* Digests of nq id lists: query q's counts[q] ids at ids + off[q] (the
 * offsets are relative to ids); ascending[q] = 1 when strictly ascending. */
void or_ids_digest(const uint32_t *ids, const uint64_t *off, const uint32_t *counts,
                   uint64_t nq, int nthreads, uint64_t *out, uint8_t *ascending);

/* CalculateStats (test/adsfilter_q_all.cu:258-307): out = mean, median,
 * stddev (population), min, max, p50, p90, p95, p99. Sorts data in place. */
void or_stats(double *data, uint64_t n, double *out9);

/* ---- text formats --------------------------------------------------------- */
/* ReadCOO (test/adsfilter_q_all.cu:127-161).  Two-call: with keys==NULL only
 * counts.  Returns pair count or -1 on open failure. */
int64_t or_read_coo(const char *path, uint32_t *keys, uint32_t *vals,
                    uint32_t *max_row, uint32_t *max_col);
/* ReadQuery (test/adsfilter_q_all.cu:164-193). */
int64_t or_read_query(const char *path, uint32_t *keys);

/* ---- MLP cost model (float64 reference for gradient checks) ---------------- */
/* Layout of params: W1[H1][F], b1[H1], W2[H2][H1], b2[H2], W3[O][H2], b3[O],
 * ReLU after layers 1 and 2; loss = mean over batch and outputs of
 * 0.5*(y_hat - y)^2. */
void or_mlp_forward(const double *params, int F, int H1, int H2, int O,
                    const double *x, int n, double *y);
double or_mlp_loss_grad(const double *params, int F, int H1, int H2, int O,
                        const double *x, const double *y, int n, double *grad);
void or_adam_step(double *params, double *m, double *v, const double *grad,
                  int np, double lr, double b1, double b2, double eps, int step);

#ifdef __cplusplus
}
#endif
#endif
