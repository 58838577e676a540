/*
 * adabool.h -- C ABI of the MI355X-native AdaBool OR-filtering engine.
 *
 * Given a sparse Boolean matrix of uint32 (key, value) pairs -- key = term,
 * value = row id -- and a query set of keys, the engine returns the bitmap and
 * the ascending id list of every value that some query key maps to:
 *     R(q) = { v | (k, v) in filter, k in q, k < nkeys }.
 *
 * Conventions
 *   - Every function returns ab_status; ab_last_error() gives a thread-local
 *     message for the last failure.  No C++ exception crosses this boundary.
 *   - Index handles own their device memory.  Contexts borrow the index, which
 *     must outlive them.  Inputs are borrowed for the duration of the call.
 *   - Pointers named d_* are device pointers (hipMalloc'd memory of the
 *     current device), h_* are host pointers.
 *   - Result bitmaps: bit v of a query's bitmap is bit (v & 31) of uint32
 *     word v >> 5 (equivalently bit v & 7 of byte v >> 3).  One query's bitmap
 *     spans ab_result_stride_bytes() = align_up(ceil(nvalues / 8), 16) bytes;
 *     padding bits are zero.  Every call WRITES the full bitmap (it is never
 *     OR-ed into caller memory), so callers need no memset.
 *   - Ids are ascending uint32 values of R(q).
 *   - Work is enqueued on the context's stream; call ab_ctx_sync() (or
 *     synchronize the stream) before reading results.
 *   - One context per host thread; an index may be shared read-only by many
 *     contexts.
 */
#ifndef ADABOOL_H
#define ADABOOL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ab_status {
    AB_OK = 0,
    AB_INVALID_ARG = 1,
    AB_OOM = 2,
    AB_DEVICE_ERROR = 3,
    AB_UNSUPPORTED = 4,
    AB_IO_ERROR = 5
} ab_status;

/* Which method evaluates a query. */
typedef enum ab_path {
    AB_PATH_AUTO = 0,      /* learned cost model if one is loaded, else threshold */
    AB_PATH_DENSE = 1,     /* dense bitmap-probe over the value-tiled key stream */
    AB_PATH_SPARSE = 2,    /* sparse load-balanced list union */
    AB_PATH_THRESHOLD = 3, /* dense iff M(q) >= tau * R */
    AB_PATH_LEARNED = 4,   /* MLP cost model; falls back to threshold without one */
    AB_PATH_IDS = 5        /* id-only output (ab_query_batch_ids); reported, never requested */
} ab_path;

/* Record types of the compressed lists (src/AdaBool/inverted_list.h:16-26). */
enum {
    AB_ARRAY1 = 0,
    AB_ARRAY4 = 1,
    AB_ARRAY8 = 2,
    AB_ARRAY16 = 3,
    AB_ARRAY32 = 4,
    AB_BITSET = 5,
    AB_NTYPES = 6
};

#define AB_NFEATURES 12

typedef struct ab_index ab_index;
typedef struct ab_ctx ab_ctx;

typedef struct ab_index_info {
    uint32_t nkeys;            /* key space: keys >= nkeys match nothing */
    uint32_t nvalues;          /* value space: bits per result bitmap */
    uint64_t nnz;              /* distinct (key, value) pairs */
    uint64_t nrec[AB_NTYPES];  /* records per type */
    uint64_t nrec_total;       /* R */
    uint64_t list_bytes;       /* sum over records of 4 + payload bytes */
    uint32_t tile_shift;       /* value tile = 2^tile_shift values */
    uint32_t ntiles;
    uint64_t nblocks;          /* key-stream blocks of the value-tiled layout */
    uint64_t key_stream_bytes; /* delta bytes of the value-tiled key stream */
    uint64_t device_bytes;     /* device memory held by the index */
    uint64_t stride_bytes;     /* result bitmap bytes per query */
} ab_index_info;

/* ---- errors --------------------------------------------------------------- */
const char *ab_last_error(void);
const char *ab_status_str(int status);

/* ---- index ---------------------------------------------------------------- */
/* Builds the compressed index from nnz (key, value) pairs and uploads it.
 * Replaces ivf_list<256,64>::BuildFrom (src/AdaBool/inverted_list.h:348-374).
 * nkeys / nvalues: 0 derives max+1 from the pairs (test/adsfilter_q_all.cu:
 * 359-360); a larger value widens the space; a smaller one is AB_INVALID_ARG.
 * Duplicate pairs are removed.  tile_shift (8..16, 0 = auto) sets the value
 * tile of the dense layout. */
ab_status ab_index_build(const uint32_t *h_keys, const uint32_t *h_values, uint64_t nnz,
                         uint32_t nkeys, uint32_t nvalues, uint32_t tile_shift,
                         ab_index **out);

/* The same build on the device, from device-resident pairs (d_keys, d_values),
 * on `stream` (NULL: the null stream): radix sort, record grouping, both
 * layouts and the key stream built by kernels, byte-identical to
 * ab_index_build's layout.  Replaces BuildFrom's std::map build
 * (src/AdaBool/inverted_list.h:348-374) and its per-array cudaMalloc+copies. */
ab_status ab_index_build_device(const uint32_t *d_keys, const uint32_t *d_values, uint64_t nnz,
                                uint32_t nkeys, uint32_t nvalues, uint32_t tile_shift, void *stream,
                                ab_index **out);

/* ReadCOO + BuildFrom (test/adsfilter_q_all.cu:127-161, 340-357): text file of
 * "key value" lines; blank lines and lines starting with '#' or '%' skipped. */
ab_status ab_index_load_coo(const char *path, uint32_t tile_shift, ab_index **out);

/* Same build without any device allocation: only info, export, decode and
 * host feature calls accept such a handle.  For hosts without a GPU. */
ab_status ab_index_build_host(const uint32_t *h_keys, const uint32_t *h_values, uint64_t nnz,
                              uint32_t nkeys, uint32_t nvalues, uint32_t tile_shift,
                              ab_index **out);

/* Binary index file: both resident layouts as built (a header, every array,
 * a 64-bit FNV-1a checksum).  Loading skips the build; host_only != 0 loads
 * without any device allocation (as ab_index_build_host).  A truncated,
 * corrupt or other-version file is AB_IO_ERROR.  The reference rebuilt its
 * index from the COO text on every run (test/adsfilter_q_all.cu:340-357). */
ab_status ab_index_save(const ab_index *ix, const char *path);
ab_status ab_index_load(const char *path, int host_only, ab_index **out);

ab_status ab_index_free(ab_index *ix);
ab_status ab_index_get_info(const ab_index *ix, ab_index_info *out);
uint64_t ab_result_stride_bytes(const ab_index *ix);

/* Copies one record type's arrays back to the host (headers: nrec u32,
 * payload: nrec * payload-bytes, segs: nkeys+1 u32 prefix counts per key, as
 * key_segments_ in src/AdaBool/inverted_list.h:461-468).  NULL pointers skip. */
ab_status ab_index_export_type(const ab_index *ix, int type, uint32_t *h_headers,
                               uint8_t *h_payload, uint32_t *h_segs);

/* Decodes every record of the value-tiled layout (the dense path's copy) back
 * to (key, value) pairs, in layout order; returns the pair count (outputs may
 * be NULL to count only, else cap pairs are written at most). */
int64_t ab_index_decode_tiles(const ab_index *ix, uint32_t *h_keys, uint32_t *h_values,
                              uint64_t cap);

/* Key-block metadata of a sorted key array, as ComputeKeyBlockMeta
 * (src/AdaBool/inverted_list.h:90-148).  Returns the padded buffer bytes; with
 * non-NULL outputs fills starts[nblocks+1], pos[nblocks+1], buf. */
uint64_t ab_key_block_meta(uint32_t step, uint64_t len, const uint32_t *keys,
                           uint32_t *starts, uint32_t *pos, uint8_t *buf);

/* ---- contexts ------------------------------------------------------------- */
/* stream: a hipStream_t, or NULL for a stream owned by the context. */
ab_status ab_ctx_create(ab_index *ix, void *stream, ab_ctx **out);
ab_status ab_ctx_free(ab_ctx *ctx);
ab_status ab_ctx_sync(ab_ctx *ctx);
void *ab_ctx_stream(ab_ctx *ctx);
/* Threshold rule: dense iff M(q) >= tau * R (M = records the query matches). */
ab_status ab_ctx_set_threshold(ab_ctx *ctx, double tau);
double ab_ctx_get_threshold(const ab_ctx *ctx);
/* The same rule inside batches of two or more queries, where one dense pass
 * over the index serves a whole group of dense queries (default 0.001). */
ab_status ab_ctx_set_batch_threshold(ab_ctx *ctx, double tau);
double ab_ctx_get_batch_threshold(const ab_ctx *ctx);
/* Dense group width (queries per key-stream pass, 1..256; host-dispatched
 * batches, mode 0, use at most 64; a device-dispatched batch of fewer
 * queries uses its own query count as the width). */
ab_status ab_ctx_set_dense_group(ab_ctx *ctx, uint32_t width);
/* Work-group size of the dense kernel (64, 128, 256 or 512 threads; default 512). */
ab_status ab_ctx_set_dense_threads(ab_ctx *ctx, uint32_t threads);
/* Per-value query-set accumulation in dense groups (on by default; used when
 * it fits the work-group's LDS, and then by every matched record).  Off: every
 * match is OR-ed into each matching query's row tile. */
ab_status ab_ctx_set_dense_columns(ab_ctx *ctx, int on);
/* Records per work item of the sparse kernel (power of two, 256..8192). */
ab_status ab_ctx_set_sparse_chunk(ab_ctx *ctx, uint32_t items);

/* Launch settings by batch size: every ab_query_batch / ab_query_batch_packed
 * call of nq queries first applies the entry with the largest min_batch <= nq
 * (the four settings above; none below the first entry's min_batch).  Entries
 * in strictly ascending min_batch (>= 1); n = 0 removes the table.  A serving
 * loop whose batches range from a few queries to its maximum uses one table
 * instead of settings tuned for one size (ab_ctx_autotune at each size gives
 * the entries; the autotuner itself ignores the table).  No reference
 * counterpart (the reference's launch parameters are compile-time constants). */
typedef struct ab_tuning {
    uint32_t min_batch;
    uint32_t group;          /* ab_ctx_set_dense_group */
    uint32_t dense_threads;  /* ab_ctx_set_dense_threads */
    uint32_t sparse_chunk;   /* ab_ctx_set_sparse_chunk */
    int32_t dense_columns;   /* ab_ctx_set_dense_columns */
} ab_tuning;
ab_status ab_ctx_set_batch_tuning(ab_ctx *ctx, const ab_tuning *entries, uint32_t n);
/* The launch settings in force (those the last batch applied, or as set);
 * out->min_batch = 0. */
ab_status ab_ctx_get_batch_settings(const ab_ctx *ctx, ab_tuning *out);

/* How ab_query_batch runs a batch: 0 host dispatch (features and decisions
 * copied back, one launch per dense group), 1 device dispatch (decisions,
 * grouping and masks on the device, then ONE persistent launch that takes
 * dense (group, tile) items and sparse chunks from a shared work counter; no
 * host round trip), 2 (default) = 1 captured once per batch shape and buffer
 * set into a hipGraph and replayed.  Per-kernel timing (ab_ctx_set_timing)
 * runs mode 2 batches as mode 1. */
ab_status ab_ctx_set_batch_mode(ab_ctx *ctx, int mode);

/* Autotuner.  Times ab_query_batch on the given batch (device keys, host
 * offsets; union only, no compaction) for every candidate of the grid
 *   dense group width {8,16,32,48,64,96,128,192,256} x dense threads {256,512}
 *   x sparse chunk {4096,8192} x dense columns {on,off},
 * `repeats` calls each after one warm-up call, and leaves the context set to
 * the fastest valid one.  Each warm-up call's result bitmaps are compared with
 * those of the settings the context held on entry (2 x nq bitmaps of device
 * scratch, freed on return); a candidate that differs is reported with its
 * mismatched word count and never chosen, and if none matches the entry
 * settings are kept and AB_DEVICE_ERROR returned.  out (may be NULL) receives
 * up to cap candidates, valid ones first, each part fastest first, with their
 * mean ms per batch; n_out the number timed.  The value
 * tile is an index layout parameter (ab_index_build's tile_shift): tune it
 * by building one index per candidate.  No reference counterpart (its launch
 * shapes are compile-time template arguments). */
typedef struct ab_tune_config {
    uint32_t group;          /* effective dense group width */
    uint32_t dense_threads;
    uint32_t sparse_chunk;
    int dense_columns;       /* effective (on only where the LDS fits) */
    double ms;               /* mean device ms per batch */
    uint64_t mismatched_words; /* result-bitmap words that differ from the entry settings' (0: valid) */
} ab_tune_config;
ab_status ab_ctx_autotune(ab_ctx *ctx, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq, int path,
                          uint32_t repeats, ab_tune_config *out, uint32_t cap, uint32_t *n_out);

/* Kernel classes reported by ab_ctx_timing. */
enum {
    AB_K_DECIDE = 0,      /* device features + dispatch decision */
    AB_K_SPARSE_ZERO = 1, /* zero-fill of sparse queries' bitmaps */
    AB_K_SPARSE_PLAN = 2, /* per-key extents + prefix scan */
    AB_K_SPARSE_UNION = 3,
    AB_K_DENSE_MASK = 4,  /* query mask set + clear */
    AB_K_DENSE_TILES = 5,
    AB_K_POPCOUNT = 6,
    AB_K_WRITE_IDS = 7,
    AB_K_FUSED_UNION = 8, /* persistent dense + sparse union (device-dispatched batches) */
    AB_K_IDS = 9,         /* id-only batch: expansion, sort, unique, write (ab_query_batch_ids) */
    AB_K_NCLASSES = 10
};
/* Brackets every kernel class of later calls with HIP events on the context
 * stream (off by default; on costs a few microseconds per call). */
ab_status ab_ctx_set_timing(ab_ctx *ctx, int on);
/* Synchronizes, then adds the elapsed ms and launch counts per class since
 * the last reset into ms[AB_K_NCLASSES] and launches[AB_K_NCLASSES]
 * (overwritten); reset != 0 clears the accumulators. */
ab_status ab_ctx_timing(ab_ctx *ctx, double *ms, uint64_t *launches, int reset);

/* ---- single query --------------------------------------------------------- */
/* One query of n host keys (duplicates and keys >= nkeys allowed).  Writes
 * the bitmap (d_bitmap, ab_result_stride_bytes() bytes, may be NULL for ids
 * only), the ascending ids (d_ids,
 * capacity nvalues, may be NULL) and the id count (d_count, device u32, may be
 * NULL).  h_chosen (may be NULL) receives the path that ran.
 * Replaces Prepare(keys) / Prepare(len, bitmap) + ListUnion_sparse /
 * ListUnion_dense (src/AdaBool/inverted_list.h:481-543,
 * test/adsfilter_q_all.cu:26-105) with the dispatch those callers lacked. */
ab_status ab_query(ab_ctx *ctx, const uint32_t *h_keys, uint32_t n, int path,
                   uint32_t *d_bitmap, uint32_t *d_ids, uint32_t *d_count, int *h_chosen);

/* Sparse method on host keys into d_res, a device buffer of res_len_bytes
 * bytes, at least ceil(nvalues / 8) (the size the reference's callers
 * allocate, test/adsfilter_q_all.cu:361-368; smaller is AB_INVALID_ARG).
 * min(res_len_bytes, ab_result_stride_bytes()) bytes are written, never more:
 * Prepare(keys) + ListUnion_sparse (src/AdaBool/inverted_list.h:481-527,
 * test/adsfilter_q_all.cu:64-105). */
ab_status ab_list_union_sparse(ab_ctx *ctx, const uint32_t *h_keys, uint32_t n,
                               uint32_t *d_res, uint64_t res_len_bytes);

/* Dense method on a host key bitmap of len_bytes bytes (bit k set = key k in
 * the query) into d_res, sized and written as for ab_list_union_sparse:
 * Prepare(len, bitmap) + ListUnion_dense
 * (src/AdaBool/inverted_list.h:530-543, test/adsfilter_q_all.cu:26-61). */
ab_status ab_list_union_dense(ab_ctx *ctx, uint64_t len_bytes, const uint8_t *h_bitmap,
                              uint32_t *d_res, uint64_t res_len_bytes);

/* ab_list_union_sparse without any allocation at call time, for callers that
 * cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.  Keys
 * are device memory (d_keys, n <= 2^18 keys, else AB_INVALID_ARG); d_res
 * holds at least ab_result_stride_bytes() bytes and receives the full
 * stride.  The plan, scan and union take their scratch (5.25 MiB) from one
 * fixed device array per GPU, held until the process ends; the call enqueues
 * kernels on the context stream only (no allocation, no copy, no memset, no
 * synchronization).  Calls that overlap on one GPU, from any contexts, share
 * that array and race: serialize them (one stream, or events). */
ab_status ab_list_union_sparse_noalloc(ab_ctx *ctx, const uint32_t *d_keys, uint32_t n,
                                       uint32_t *d_res, uint64_t res_len_bytes);

/* Batched dense method on B device bitmaps laid out query-major, each
 * len_bytes bytes (a multiple of 4); query i's result goes to
 * d_res + i * res_len_u32.  PrepareBatchDevice + RunArrays
 * (src/AdaBool/inverted_list.h:591-605, src/AdaBool/inverted_list.cu.h:
 * 224-228), for any B >= 1 (the reference handled 1..32). */
ab_status ab_list_union_dense_batch(ab_ctx *ctx, uint32_t batch, uint64_t len_bytes,
                                    const uint8_t *d_bitmaps, uint32_t *d_res,
                                    uint32_t res_len_u32);

/* Batched sparse method: n flattened host keys, h_qidx[i] = query of key i
 * (0..nq-1); results at d_res + q * res_len_bytes / 4.  BatchListUnion_sparse
 * (test/adsfilter_batch.cu:39-79, 473-484). */
ab_status ab_list_union_sparse_batch(ab_ctx *ctx, const uint32_t *h_keys,
                                     const uint32_t *h_qidx, uint32_t n, uint32_t nq,
                                     uint32_t *d_res, uint32_t res_len_bytes);

/* ---- batches -------------------------------------------------------------- */
/* nq queries whose keys are concatenated in device memory d_keys, query i
 * owning d_keys[h_qoff[i] .. h_qoff[i+1]).  path applies to every query
 * (AUTO/THRESHOLD/LEARNED dispatch per query).  Outputs (each may be NULL):
 *   d_bitmaps  nq * stride bytes (query i at i * stride),
 *   d_ids      query i's ids at d_ids + i * ids_stride (ids_stride >= the
 *              query's count; nvalues is always enough),
 *   d_counts   nq u32 id counts,
 *   h_paths    nq bytes: the path each query took (AB_PATH_DENSE/SPARSE).
 * Queries of either path are packed into one union launch of the persistent
 * kernel; the whole batch is replayed from a cached HIP graph when the same
 * shape recurs.  Batches of up to 4 queries with at most 65536 keys choose
 * their paths on the host instead: their keys are copied back and the call
 * waits for that copy. */
ab_status ab_query_batch(ab_ctx *ctx, const uint32_t *d_keys, const uint64_t *h_qoff,
                         uint32_t nq, int path, uint32_t *d_bitmaps, uint32_t *d_ids,
                         uint64_t ids_stride, uint32_t *d_counts, uint8_t *h_paths);

/* ab_query_batch with ids packed at caller-chosen offsets: query i's ids go
 * to d_ids + h_id_off[i] (u32 units; ranges must not overlap, and offsets
 * that are multiples of 4 take the 128-bit store path).  Each range must hold
 * the query's count, which is at most its id bound (ab_query_id_bound); the
 * library cannot see the size of a device buffer, and a shorter range is
 * written past.  Variable-size id
 * lists packed on the device, where the reference's batch caller reads every
 * bitmap back and collects each query's ids on the host
 * (test/adsfilter_batch.cu:505-537). */
ab_status ab_query_batch_packed(ab_ctx *ctx, const uint32_t *d_keys, const uint64_t *h_qoff,
                                uint32_t nq, int path, uint32_t *d_bitmaps, uint32_t *d_ids,
                                const uint64_t *h_id_off, uint32_t *d_counts, uint8_t *h_paths);

/* Id-only batch: query i's ascending ids at d_ids + h_id_off[i] (u32 units)
 * and its count at d_counts[i] (NULL: not returned), formed without any
 * result bitmap of nvalues bits in device memory.  Two methods:
 *   by sort: the records are expanded into values, radix-sorted and
 *   deduplicated on the device, so memory and traffic follow the values the
 *   lists hold, whatever nvalues (up to 2^32 - 1);
 *   by value slice (nvalues <= 2^28 and every record type below 2^29
 *   records): the records are routed by slice of 2^18 values and each
 *   (query, slice) is OR-ed into a 32 KB LDS bitmap and expanded straight to
 *   ids; its cost follows nq x slices, so it serves small batches and dense
 *   results.
 * The automatic choice (ab_ctx_set_id_method) takes the slices for batches
 * of at most 8192 (query, slice) items or whose values fill at least 1/20 of
 * the slices' value space, else the sort.  It serves indices whose per-query
 * bitmaps would not fit and queries whose results are a small part of
 * nvalues.  Query i needs room for ab_query_id_bound ids; ranges must not
 * overlap.  The batch runs in parts of at most the context's
 * id budget of expanded values (ab_ctx_set_id_budget; a larger query forms a
 * part of its own), with one host round trip per call.  The result equals
 * ab_query_batch's ids for every query.  No reference counterpart: the
 * reference always forms a bitmap of nvalues bits per query and collects ids
 * from it on the host (test/adsfilter_q_all.cu:108-119). */
ab_status ab_query_batch_ids(ab_ctx *ctx, const uint32_t *d_keys, const uint64_t *h_qoff,
                             uint32_t nq, uint32_t *d_ids, const uint64_t *h_id_off,
                             uint32_t *d_counts);
/* Upper bound of a query's id count from host keys: the sum of its in-range
 * keys' list lengths (repeats included), capped at nvalues. */
ab_status ab_query_id_bound(const ab_index *ix, const uint32_t *h_keys, uint32_t n,
                            uint64_t *bound);
/* Expanded values per part of ab_query_batch_ids by sort (>= 4096; default
 * 2^27, 3 GiB of scratch); the slice method takes parts of at most 4x as many
 * records (4-byte references), and at most 2^29. */
ab_status ab_ctx_set_id_budget(ab_ctx *ctx, uint64_t values);
/* Method of ab_query_batch_ids: 0 (default) automatic; 1 always by sort;
 * 2 by value slice wherever it applies (else by sort). */
ab_status ab_ctx_set_id_method(ab_ctx *ctx, int method);

/* Bitmap -> ascending ids for nq bitmaps at d_bitmaps + i * stride
 * (CollectIdsFromBitmap, test/adsfilter_q_all.cu:108-119, on the device). */
ab_status ab_compact_ids(ab_ctx *ctx, const uint32_t *d_bitmaps, uint32_t nq,
                         uint32_t *d_ids, uint64_t ids_stride, uint32_t *d_counts);

/* ---- cost model ----------------------------------------------------------- */
/* Per-query features (AB_NFEATURES floats per query) computed on the device
 * for a device batch; h_features receives nq * AB_NFEATURES floats and h_m
 * (may be NULL) the matched-record counts M(q). */
ab_status ab_query_features(ab_ctx *ctx, const uint32_t *d_keys, const uint64_t *h_qoff,
                            uint32_t nq, float *h_features, uint64_t *h_m);
/* Same features from host keys, computed on the host. */
ab_status ab_query_features_host(const ab_index *ix, const uint32_t *h_keys, uint32_t n,
                                 float *h_feature, uint64_t *h_m);

/* ---- cost model (FP32 MLP 12 -> 64 -> 64 -> 2, hand-written HIP kernels) --
 * Predicts log execution time of the dense (output 0) and sparse (output 1)
 * path from AB_NFEATURES standardized features; the dispatcher takes the
 * smaller.  Parameters: W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64],
 * b3[2] (row-major, ReLU after the hidden layers).  A context holds two
 * models: AB_MODEL_SINGLE prices one query run on its own (ab_query),
 * AB_MODEL_BATCH prices a query's share of a batch (ab_query_batch with two
 * or more queries, where one dense pass serves a whole group).
 * AB_PATH_AUTO / AB_PATH_LEARNED use the model of the call's kind when it is
 * loaded and fall back to the threshold rule otherwise.  No reference
 * counterpart: the reference runs both paths per query and has no
 * dispatcher. */
enum { AB_MODEL_SINGLE = 0, AB_MODEL_BATCH = 1 };
uint32_t ab_mlp_nparams(void);
ab_status ab_mlp_set(ab_ctx *ctx, int slot, const float *params, const float *mean, const float *stdv);
ab_status ab_mlp_get(ab_ctx *ctx, int slot, float *params, float *mean, float *stdv);
ab_status ab_mlp_clear(ab_ctx *ctx, int slot);
/* Device inference for n host feature rows: h_y gets n*2 predictions, h_sel
 * (may be NULL) the chosen paths. */
ab_status ab_mlp_predict(ab_ctx *ctx, int slot, const float *h_x, uint32_t n, float *h_y,
                         uint8_t *h_sel);
/* Loss gradient over n (1..4096) samples at the loaded parameters, computed
 * by the training kernel in gradient-only mode (loss = mean over samples and
 * outputs of 0.5*(y_hat - y)^2). */
ab_status ab_mlp_grad(ab_ctx *ctx, int slot, const float *h_x, const float *h_y, uint32_t n,
                      float *h_grad, float *h_loss);
/* Adam training on the GPU over n samples (targets n*2 log times), shuffled
 * minibatches of `batch` (<= 4096) rows for `epochs` epochs.  Starts from the
 * loaded model, or from a seeded initialization with mean/std taken from the
 * data.  loss_curve (may be NULL) receives the mean loss of each epoch. */
ab_status ab_mlp_train(ab_ctx *ctx, int slot, const float *h_x, const float *h_y, uint32_t n,
                       uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss_curve);

/* ---- serving loop ----------------------------------------------------------
 * A resident index served to requests submitted from any host thread.  Each
 * request reserves result memory for its id upper bound (sum of its keys'
 * list lengths, capped at nvalues) in a device arena at admission, in arrival
 * order; admitted requests are batched continuously into device-dispatched
 * batches (ab_query_batch_packed) on `inflight` contexts, each with its own
 * stream, so one batch's compaction overlaps the next one's union.  Results
 * stay valid until ab_server_release.  Replaces the reference's harness loop
 * over query files (test/adsfilter_batch.cu:428-568), which ran fixed
 * batches synchronously. */
typedef struct ab_server ab_server;
typedef struct ab_server_config {
    uint32_t max_batch;   /* queries per batch (1..4096), default 1024 */
    uint32_t max_wait_us; /* a partial batch leaves after this wait when the device is busy (200) */
    uint32_t inflight;    /* batches in flight, one context and stream each (1..8, default 2) */
    uint64_t arena_bytes; /* device memory for result ids (default 4 GiB) */
    int path;             /* dispatch of every request (default AB_PATH_AUTO) */
    int deliver;          /* 0: ids in device memory; 1: ids copied to a pinned host mirror */
    int batch_mode;       /* ab_ctx_set_batch_mode of the contexts (default 1) */
    int ids_only;         /* 1: batches run ab_query_batch_ids (no bitmaps; path AB_PATH_IDS) */
} ab_server_config;
typedef struct ab_result {
    uint64_t ticket;
    int status;           /* AB_OK, or AB_OOM when the request can never fit the arena */
    int path;             /* path taken (AB_PATH_DENSE / AB_PATH_SPARSE / AB_PATH_IDS) */
    uint32_t count;       /* ids */
    const uint32_t *ids;  /* device (deliver 0) or host (deliver 1) pointer, valid until release */
    double latency_ms;    /* submit -> results ready */
} ab_result;
typedef struct ab_server_stats_t {
    uint64_t completed, batches, dense_queries;
    double mean_batch, qps;
    double lat_mean_ms, lat_p50_ms, lat_p90_ms, lat_p99_ms, lat_max_ms;
    uint64_t arena_used_bytes;
} ab_server_stats_t;
ab_status ab_server_config_default(ab_server_config *cfg);
ab_status ab_server_create(ab_index *ix, const ab_server_config *cfg, ab_server **out);
/* Context i (< inflight) of the server, for settings (thresholds, cost
 * models, dense group...) made before the first submission. */
ab_ctx *ab_server_ctx(ab_server *srv, uint32_t i);
/* Thread-safe; the keys are copied. */
ab_status ab_server_submit(ab_server *srv, const uint32_t *h_keys, uint32_t n, uint64_t *ticket);
/* Up to cap completed requests in completion order; waits up to wait_ms
 * (-1: until one completes, 0: no wait) when none is ready. */
ab_status ab_server_poll(ab_server *srv, ab_result *out, uint32_t cap, uint32_t *n, int wait_ms);
/* Returns a completed request's result memory to the arena. */
ab_status ab_server_release(ab_server *srv, uint64_t ticket);
/* Latency percentiles (CalculateStats definition) and throughput since the
 * last reset, over requests returned by ab_server_poll. */
ab_status ab_server_stats(ab_server *srv, ab_server_stats_t *out, int reset);
/* Stops after finishing every submitted request; frees all memory. */
ab_status ab_server_free(ab_server *srv);
/* Open-loop load: query i of (h_keys, h_qoff) arrives at Poisson times of
 * rate_qps (0: all at once); results are collected and released as they
 * complete.  h_counts / h_lat_ms (may be NULL) get each query's id count and
 * latency from its scheduled arrival; wall_s the time to the last result. */
ab_status ab_server_replay(ab_server *srv, const uint32_t *h_keys, const uint64_t *h_qoff, uint32_t nq,
                           double rate_qps, uint64_t seed, uint32_t *h_counts, double *h_lat_ms,
                           double *wall_s);

/* ---- synthetic data ------------------------------------------------------- */
typedef struct ab_dataset ab_dataset;
/* Filter pairs: nvalues rows, each drawing a row length L then L distinct
 * keys from Zipf(s) over nkeys keys (ranks mapped to keys by a permutation
 * seeded with perm_seed).  len_dist 0: L ~ U{lo..hi}; 1: L = round(
 * LogNormal(ln p1, p2)) clipped to [lo, hi]. */
ab_status ab_gen_pairs(uint64_t seed, uint64_t perm_seed, uint32_t nvalues, uint32_t nkeys,
                       int len_dist, double p1, double p2, uint32_t lo, uint32_t hi,
                       double zipf_s, int nthreads, ab_dataset **out);
/* Queries: size_dist 0: n = lo; 1: round(LogNormal(ln p1, p2)) clipped to
 * [lo, hi]; 2: discrete power law P(n) ~ n^-p1 on [lo, hi].  key_dist 0:
 * distinct uniform keys; 1: distinct Zipf(s)-weighted keys (same permutation
 * as the pairs for the same perm_seed). */
ab_status ab_gen_queries(uint64_t seed, uint64_t perm_seed, uint32_t nkeys, uint32_t nq,
                         int size_dist, double p1, double p2, uint32_t lo, uint32_t hi,
                         int key_dist, double zipf_s, ab_dataset **out);
/* Pairs: a = keys, b = values, len = nnz.  Queries: a = keys (concatenated),
 * offsets = nq+1 prefix, len = total keys. */
uint64_t ab_dataset_len(const ab_dataset *d);
uint64_t ab_dataset_count(const ab_dataset *d);
const uint32_t *ab_dataset_a(const ab_dataset *d);
const uint32_t *ab_dataset_b(const ab_dataset *d);
const uint64_t *ab_dataset_offsets(const ab_dataset *d);
void ab_dataset_free(ab_dataset *d);

/* ---- text loaders --------------------------------------------------------- */
/* ReadQuery (test/adsfilter_q_all.cu:164-193): returns the key count (or -1)
 * and fills keys when non-NULL. */
int64_t ab_read_query(const char *path, uint32_t *keys, uint64_t cap);
/* FindQueryFiles (test/adsfilter_q_all.cu:211-231): the regular files of dir
 * (symlinks followed) whose whole name matches query_vec_[0-9]+\.coo, as
 * dir + "/" + name, sorted lexicographically as byte strings (so
 * query_vec_10.coo comes before query_vec_2.coo).  Returns the file count, or
 * -1 when dir cannot be read (the reference printed the error and returned no
 * files).  *need (may be NULL) gets the bytes of all paths, NUL-terminated one
 * after another; they are written to buf only when cap holds them all. */
int64_t ab_find_query_files(const char *dir, char *buf, uint64_t cap, uint64_t *need);
/* ReadCOO into caller buffers (two-call: keys == NULL counts). */
int64_t ab_read_coo(const char *path, uint32_t *keys, uint32_t *values, uint64_t cap,
                    uint32_t *max_key, uint32_t *max_value);

/* ---- statistics ----------------------------------------------------------- */
/* CalculateStats (test/adsfilter_q_all.cu:258-307): out = mean, median,
 * stddev, min, max, p50, p90, p95, p99 (sorts data in place). */
void ab_stats(double *data, uint64_t n, double *out9);

#ifdef __cplusplus
}
#endif
#endif /* ADABOOL_H */
